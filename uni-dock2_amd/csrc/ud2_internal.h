// Internal structures shared by the host API (ud2_api.cpp) and the kernels
// (ud2_kernels.hip). Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ud2 {

constexpr int kWave = 64;
constexpr int kMaxApl = 4;                  // atoms per lane: natom <= 256
constexpr int kMaxAtoms = kWave * kMaxApl;  // pair records pack atom indices in 8 bits
constexpr int kMaxTors = 57;                // 7 + ntors <= 64: one pose-vector / BFGS component per lane
constexpr int kWavesPerBlock = 4;
constexpr int kNumTypes = 21;

// One ligand of a ligand set on the device.
struct DevLig {
    int32_t natom, ntors, npair, nheavy;
    int32_t atom_off;   // into atom_x0 (atoms)
    int32_t tors_off;   // into tors_* (torsions)
    int32_t pair_off;   // into pairs
    int32_t apl;        // ceil(natom / 64)
};

// Device view of a ligand set.
struct LigSetDev {
    const DevLig* ligs;
    const float4* atom_x0;                 // x, y, z, type (int bits)
    const int2* tors_axis;                 // (from, to)
    const float* tors_dihe0;               // input dihedral (radians)
    const int2* tors_range;                // (global offset into ranges, count)
    const float2* ranges;                  // (lo, hi) radians
    const unsigned long long* tors_mask;   // [ntors_total][kMaxApl] rotated-atom bit masks
    const uint2* pairs;                    // {i | j << 8 | hyd << 16 | hbond << 17, bits of R_i + R_j}
};

// Device view of a scene: receptor, box and (optional) grid maps.
struct SceneDev {
    const float4* rec;  // heavy receptor atoms: x, y, z, type (int bits)
    int32_t nrec;
    float box_lo[3], box_hi[3];
    const float* maps;  // [slot][nz][ny][nx]
    int32_t gn[3];
    float gorigin[3];
    float gspacing, ginv;
    const int32_t* slot_tab;  // [kNumTypes] map of each XS type (-1 = none), device memory
};

// Pose buffers of a ligand set with nchain poses per ligand, laid out by
// ligand: x at (atom_off * nchain + chain * natom) * 3, dihedrals at
// tors_off * nchain + chain * ntors, centre / rot_vec at (lig * nchain + chain) * 3,
// energy at lig * nchain + chain.
struct PoseBufs {
    float* x;
    float* c;
    float* rv;
    float* dihe;
    float* e;
    int32_t nchain;
};

enum InterMode : int32_t { kDirect = 0, kGrid = 1 };

struct McArgs {
    LigSetDev L;
    SceneDev S;
    PoseBufs P;
    const int32_t* order;  // ligands in the work order of this launch
    int32_t n_order;
    int32_t* counter;      // work-queue head, zero at launch
    int32_t mode;
    int32_t mc_steps, opt_steps, randomize;
    uint32_t seed;
    int32_t lig_index_base;
    int32_t lds_floats_per_wave;
    int32_t hdim_ld;       // leading dimension of the per-wave Hessian (odd)
    unsigned long long* stats;  // optional: evals, gather bytes, pair evals, chains
};

struct OptArgs {
    LigSetDev L;
    SceneDev S;
    PoseBufs P;
    const int32_t* sel;  // selected poses: lig * nchain + chain
    const int32_t* pos;  // output slot of each selected pose (null: its index in sel)
    int32_t n_sel;
    int32_t mode;
    int32_t max_steps;
    float v_cap, slope;
    int32_t lds_floats_per_wave;
    int32_t hdim_ld;
    float* out_e;   // energy of each selected pose (may be null)
    float* out_g;   // [n_sel][64] gradient (may be null)
    const float* dir;   // [n_sel][64] update direction (pose-update hook)
    float alpha;
    int32_t op;     // 0 BFGS, 1 energy+gradient, 2 pose update, 3 line search from H = I, 4 reset
};

struct PoseVecArgs {
    LigSetDev L;
    SceneDev S;
    int32_t lig;
    int64_t n;
    const float* pv;   // [n][7 + ntors]
    float* e;          // [n]
    float* g;          // [n][6 + ntors]
    float v_cap, slope;
    int32_t mode;
    int32_t lds_floats_per_wave;
};

struct ClusterArgs {
    LigSetDev L;
    const float* x;
    const float* e;
    int32_t nchain;
    int32_t n_lig;
    float limit;
    int32_t* leaders;   // [n_lig][nchain]
    int32_t* nleaders;  // [n_lig]
};

struct ScoreArgs {
    LigSetDev L;
    SceneDev S;
    const float* x;
    int32_t nchain;
    const int32_t* sel;
    int32_t n_sel;
    float* out3;  // [n_sel][3]: intra, inter, penalty
};

struct GridArgs {
    SceneDev S;
    const int32_t* probe_types;  // [nslot]
    int32_t nslot;
    float* maps;
};

// Launchers (ud2_kernels.hip); each returns a hipError_t.
hipError_t launch_grid_build(const GridArgs& a, hipStream_t s);
hipError_t launch_mc(McArgs a, int apl, int n_blocks, hipStream_t s);
hipError_t launch_opt(OptArgs a, int apl, hipStream_t s);
hipError_t launch_posevec(PoseVecArgs a, int apl, int n_blocks, hipStream_t s);
hipError_t launch_cluster(const ClusterArgs& a, hipStream_t s);
hipError_t launch_final_score(const ScoreArgs& a, int apl, hipStream_t s);
hipError_t launch_fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s);
int mc_blocks_per_cu(int apl, int lds_bytes_per_block);
// LDS floats of a wave slice before its Hessian tile.
// coordinates + pair forces (7 * 64 * apl) and two parked poses (64 * (3 * apl + 1) + 8 each)
__host__ __device__ constexpr int lds_fixed_floats(int apl) { return 7 * kWave * apl + 2 * (kWave * (3 * apl + 1) + 8); }
int posevec_blocks_per_cu(int apl, int lds_bytes_per_block);

}  // namespace ud2
