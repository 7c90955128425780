// Round kernels of the learned-proposal Monte Carlo search (no reference
// counterpart; the plain search is mc_kernel, after E/src/cuda/mc.cu:275-398).
//
// One MC step of every chain is one round of launches:
//   lmc_propose_kernel  16 candidate mutations of the accepted pose, each
//                       described by 80 pooled grid-map features (bf16);
//   mlp_fwd_kernel      the network scores the 16 candidates of every chain;
//   lmc_step_kernel     the top_k candidates by score are rebuilt and scored
//                       with the exact search energy, the best is BFGS-refined
//                       and goes through the reference's Metropolis test.
// Policies: plain (one random candidate, no network: the reference's step),
// exact (n_cand candidates all scored exactly, no network), learned (above),
// collect (every candidate BFGS-refined; its energy is the training label;
// the chain moves on with candidate 0, i.e. the plain trajectory).
// A candidate is identified by the Philox block where its mutation starts, so
// rebuilding it replays the same draws. Every chain's state (accepted pose,
// best pose, next Philox block) lives in HBM between rounds, so the rounds of
// a search can be captured in one HIP graph.
#include "ud2_device.h"

namespace ud2 {

// Feature class of each XS type (E/src/constants/constants.h:48-68); -1 = hydrogen.
static __constant__ int c_fclass[kNumTypes] = {-1, 1, 0, 1, 2, 3, 4, 3, 5, 6, 5, 6, 0, 1, 7, 7, 0, 0, 0, 0, 7};

__device__ __forceinline__ uint16_t f2bf_l(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) {
        if (u & 0x007FFFFFu) u |= 0x00400000u;
        return uint16_t(u >> 16);
    }
    u += 0x7FFFu + ((u >> 16) & 1u);
    return uint16_t(u >> 16);
}

// Features of pose P (80 bf16 into xrow). Per heavy atom: its own map's value e
// and gradient g at its position, statistics [1, min(e,0), clamp(e,0,1),
// log1p(max(e,0)), e < -0.3, min(box overshoot,4), min(|g|,8)(e<=0),
// min(|g|,8)(e>0)] pooled by feature class into xrow[8 class + stat]; move
// globals in xrow[64..79]. The per-atom rows go through the LDS scratch at
// `sc` (8 floats per atom) and lane j sums feature j.
template <int APL>
__device__ void pose_features(const Ctx<APL>& C, const WPose<APL>& P, const Move& mv, float e_acc, int sc,
                              uint16_t* xrow) {
    const SceneDev& S = *C.S;
    const int ln = lane();
    float curl = 0.0f;
#pragma unroll
    for (int k = 0; k < APL; k++) {
        const int a = ln + kWave * k;
        float cls = -1.0f, p1 = 0.0f, p2 = 0.0f, p3 = 0.0f, p4 = 0.0f, p5 = 0.0f, p6 = 0.0f, p7 = 0.0f;
        if (heavy(C, k)) {
            const int t = C.type[k];
            const int slot = S.slot_tab[t];
            float gx = 0.0f, gy = 0.0f, gz = 0.0f, e = 0.0f;
            if (slot >= 0) e = grid_eval(S, slot, P.x[k], P.y[k], P.z[k], gx, gy, gz);
            const float over = fmaxf(S.box_lo[0] - P.x[k], 0.0f) + fmaxf(P.x[k] - S.box_hi[0], 0.0f) +
                               fmaxf(S.box_lo[1] - P.y[k], 0.0f) + fmaxf(P.y[k] - S.box_hi[1], 0.0f) +
                               fmaxf(S.box_lo[2] - P.z[k], 0.0f) + fmaxf(P.z[k] - S.box_hi[2], 0.0f);
            const float gc = fminf(sqrtf(gx * gx + gy * gy + gz * gz), 8.0f);
            const float ep = fmaxf(e, 0.0f);
            cls = float(c_fclass[t]);
            p1 = fminf(e, 0.0f);
            p2 = fminf(ep, 1.0f);
            p3 = log1pf(fminf(ep, 1e30f));
            p4 = e < -0.3f ? 1.0f : 0.0f;
            p5 = fminf(over, 4.0f);
            p6 = e <= 0.0f ? gc : 0.0f;
            p7 = e > 0.0f ? gc : 0.0f;
            curl += e > 0.0f ? e / (1.0f + e) : e;
        }
        if (a < C.natom) {
            float* r = s_lds + sc + 8 * a;
            r[0] = cls; r[1] = p1; r[2] = p2; r[3] = p3; r[4] = p4; r[5] = p5; r[6] = p6; r[7] = p7;
        }
    }
    wave_lds_fence();
    const float my_cls = float(ln >> 3);
    const int st = ln & 7;
    float f = 0.0f;
    for (int a = 0; a < C.natom; a++) {
        const float c = s_lds[sc + 8 * a];
        const float v = st == 0 ? 1.0f : s_lds[sc + 8 * a + st];
        f += c == my_cls ? v : 0.0f;
    }
    xrow[ln] = f2bf_l(f);
    curl = wave_sum(curl);
    if (ln < 16) {
        float g = 0.0f;
        if (ln == 0) g = mv.kind == 0 ? 1.0f : 0.0f;
        if (ln == 1) g = mv.kind == 1 ? 1.0f : 0.0f;
        if (ln == 2) g = mv.kind == 2 ? 1.0f : 0.0f;
        if (ln == 3) g = mv.mag;
        if (ln == 4) g = float(C.nheavy) / 64.0f;
        if (ln == 5) g = float(C.ntors) / 16.0f;
        if (ln == 6) g = fminf(fmaxf(e_acc, -20.0f), 20.0f) / 10.0f;
        if (ln == 7) g = mv.moved;
        if (ln == 8) g = fminf(fmaxf(curl, -20.0f), 20.0f) / 10.0f;
        xrow[kPooled + ln] = f2bf_l(g);
    }
    wave_lds_fence();
}

// Round 0: input pose, randomisation (E/src/cuda/mc.cu:109-134); the accepted
// pose starts at energy 999 as the reference's does, the best at 1e9.
template <int APL>
__global__ __launch_bounds__(kWave* kWavesPerBlock) void lmc_init_kernel(LmcArgs A) {
    const int w = threadIdx.x / kWave;
    const int base = w * A.lds_floats_per_wave;
    const int nchain = A.acc.nchain;
    const int n_items = A.n_order * nchain;
    for (;;) {
        int item = 0;
        if (lane() == 0) item = atomicAdd(A.counter, 1);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= n_items) break;
        const int lig = A.order[item / nchain];
        const int chain = item % nchain;
        Ctx<APL> C;
        load_ctx(C, A.L, A.S, lig, A.mode, base, A.hdim_ld);
        Rng rng{A.seed, uint32_t(lig + A.lig_index_base), uint32_t(chain), 0u};
        WPose<APL> p;
        initial_pose(C, A.L, lig, p);
        if (A.randomize) randomize(C, p, rng, chain, nchain);
        store_pose(A.L, A.acc, lig, chain, p);
        p.e = 1e9f;
        store_pose(A.L, A.best, lig, chain, p);
        if (lane() == 0) A.draw[(size_t)lig * nchain + chain] = rng.draw;
    }
}

// Candidate mutations of the accepted pose and their features.
template <int APL>
__global__ __launch_bounds__(kWave* kWavesPerBlock) void lmc_propose_kernel(LmcArgs A) {
    const int w = threadIdx.x / kWave;
    const int base = w * A.lds_floats_per_wave;
    const int nchain = A.acc.nchain;
    const int n_items = A.n_order * nchain;
    const int ln = lane();
    for (;;) {
        int item = 0;
        if (ln == 0) item = atomicAdd(A.counter, 1);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= n_items) break;
        const int lig = A.order[item / nchain];
        const int chain = item % nchain;
        const size_t pid = (size_t)lig * nchain + chain;
        Ctx<APL> C;
        load_ctx(C, A.L, A.S, lig, A.mode, base, A.hdim_ld);
        WPose<APL> acc;
        load_pose<APL>(A.L, A.acc, lig, chain, acc);
        save_pose(C, acc, 1);
        Rng rng{A.seed, uint32_t(lig + A.lig_index_base), uint32_t(chain), A.draw[pid]};
        uint32_t my_start = 0;
        for (int c = 0; c < A.n_cand; c++) {
            if (ln == c) my_start = rng.draw;
            WPose<APL> cand;
            restore_pose(C, cand, 1);
            const Move mv = mutate(C, cand, rng);
            pose_features(C, cand, mv, acc.e, C.lh, A.x + (pid * kGroup + c) * kFeat);
        }
        if (ln < A.n_cand) A.cand_draw[pid * kGroup + ln] = my_start;
        if (ln == 0) A.draw[pid] = rng.draw;
        if (A.stats && ln == 0)
            atomicAdd(&A.stats[1], (unsigned long long)A.n_cand * 32ull * (unsigned long long)C.nheavy);
    }
}

// Selection, BFGS and the Metropolis step (mc.cu:340-392) of one round.
template <int APL>
__global__ __launch_bounds__(kWave* kWavesPerBlock) void lmc_step_kernel(LmcArgs A) {
    const int w = threadIdx.x / kWave;
    const int base = w * A.lds_floats_per_wave;
    const int nchain = A.acc.nchain;
    const int n_items = A.n_order * nchain;
    const int ln = lane();
    for (;;) {
        int item = 0;
        if (ln == 0) item = atomicAdd(A.counter, 1);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= n_items) break;
        const int lig = A.order[item / nchain];
        const int chain = item % nchain;
        const size_t pid = (size_t)lig * nchain + chain;
        Ctx<APL> C;
        load_ctx(C, A.L, A.S, lig, A.mode, base, A.hdim_ld);
        WPose<APL> acc;
        load_pose<APL>(A.L, A.acc, lig, chain, acc);
        save_pose(C, acc, 1);
        const float e_acc = acc.e;
        float best_e = uni(A.best.e[pid]);
        Rng rng{A.seed, uint32_t(lig + A.lig_index_base), uint32_t(chain), A.draw[pid]};
        WPose<APL> cand = acc;
        if (A.policy == kPolicyPlain) {
            mutate(C, cand, rng);
            bfgs(C, cand, A.opt_steps, kVCap, kSlope);
        } else if (A.policy == kPolicyCollect) {
            // every candidate refined; candidate 0 (processed last) continues the chain
            for (int c = A.n_cand - 1; c >= 0; c--) {
                Rng r2 = rng;
                r2.draw = A.cand_draw[pid * kGroup + c];
                restore_pose(C, cand, 1);
                mutate(C, cand, r2);
                bfgs(C, cand, A.opt_steps, kVCap, kSlope);
                if (ln == 0) A.label[pid * kGroup + c] = cand.e;
            }
        } else {
            // exact: all n_cand drawn here in order; learned: the top_k by score, rebuilt
            int rank = 0;
            if (A.policy == kPolicyLearned) {
                float p = ln < A.n_cand ? A.pred[pid * kGroup + ln] : INFINITY;
                p = p == p ? p : INFINITY;
                for (int j = 0; j < A.n_cand; j++) {
                    const float pj = rdl(p, j);
                    rank += (pj < p || (pj == p && j < ln)) ? 1 : 0;
                }
            }
            // candidates are scored one at a time and the winner rebuilt from its
            // first Philox block, so only one candidate pose is live at a time
            const int n_try = A.policy == kPolicyLearned ? A.top_k : A.n_cand;
            float be = INFINITY;
            uint32_t best_start = rng.draw;
            for (int r = 0; r < n_try; r++) {
                uint32_t start = rng.draw;
                if (A.policy == kPolicyLearned) {
                    const unsigned long long m = __ballot(ln < A.n_cand && rank == r);
                    start = A.cand_draw[pid * kGroup + (__ffsll((long long)m) - 1)];
                }
                Rng r2 = rng;
                r2.draw = start;
                WPose<APL> t;
                restore_pose(C, t, 1);
                mutate(C, t, r2);
                if (A.policy != kPolicyLearned) rng.draw = r2.draw;
                Forces<APL> F;
                const float e = energy_forces(C, t, kVCap, kSlope, F);
                if (e < be || r == 0) { be = e; best_start = start; }
            }
            Rng r3 = rng;
            r3.draw = best_start;
            restore_pose(C, cand, 1);
            mutate(C, cand, r3);
            bfgs(C, cand, A.opt_steps, kVCap, kSlope);
        }
        const uint4 b = rng.next();
        const bool accepted = (cand.e < e_acc) || (u01(b.x) < __expf((e_acc - cand.e) / 1.2f));
        if (A.step == 0 || accepted) {
            if (cand.e < best_e) {
                if (A.opt_steps > 0) bfgs(C, cand, A.opt_steps, kVCapAuth, kSlope);
                if (cand.e < best_e) {
                    store_pose(A.L, A.best, lig, chain, cand);
                    best_e = cand.e;
                }
            }
            store_pose(A.L, A.acc, lig, chain, cand);
        }
        if (ln == 0) A.draw[pid] = rng.draw;
        if (A.stats && ln == 0) {
            const unsigned long long ev = (unsigned long long)C.evals;
            atomicAdd(&A.stats[0], ev);
            atomicAdd(&A.stats[1], ev * 32ull * (unsigned long long)C.nheavy);
            atomicAdd(&A.stats[2], ev * (unsigned long long)C.npair);
            if (A.step == 0) atomicAdd(&A.stats[3], 1ull);
        }
    }
}

// Features of selected poses (no move: kind -1, accepted energy 0).
template <int APL>
__global__ __launch_bounds__(kWave* kWavesPerBlock) void features_kernel(FeatArgs A) {
    const int w = threadIdx.x / kWave;
    const int s = blockIdx.x * kWavesPerBlock + w;
    if (s >= A.n_sel) return;
    const int pid = A.sel[s];
    const int lig = pid / A.P.nchain, chain = pid % A.P.nchain;
    Ctx<APL> C;
    load_ctx(C, A.L, A.S, lig, kGrid, w * A.lds_floats_per_wave, 1);
    WPose<APL> p;
    load_pose<APL>(A.L, A.P, lig, chain, p);
    Move mv;
    mv.kind = -1; mv.mag = 0.0f; mv.moved = 0.0f;
    pose_features(C, p, mv, 0.0f, C.lh, A.x + (size_t)s * kFeat);
}

// ------------------------------------------------------------- launchers
#define UD2_APL_SWITCH(apl, CALL) \
    switch (apl) {                \
        case 1: CALL(1);          \
        case 2: CALL(2);          \
        case 3: CALL(3);          \
        case 4: CALL(4);          \
        default: return hipErrorInvalidValue; \
    }

template <int APL, int K>
static hipError_t launch_lmc_t(const LmcArgs& a, int n_blocks, hipStream_t s) {
    const size_t lds = (size_t)a.lds_floats_per_wave * kWavesPerBlock * sizeof(float);
    if (K == 0) hipLaunchKernelGGL(lmc_init_kernel<APL>, dim3(n_blocks), dim3(kWave * kWavesPerBlock), lds, s, a);
    if (K == 1) hipLaunchKernelGGL(lmc_propose_kernel<APL>, dim3(n_blocks), dim3(kWave * kWavesPerBlock), lds, s, a);
    if (K == 2) hipLaunchKernelGGL(lmc_step_kernel<APL>, dim3(n_blocks), dim3(kWave * kWavesPerBlock), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_lmc_init(LmcArgs a, int apl, int n_blocks, hipStream_t s) {
#define CALL(N) return launch_lmc_t<N, 0>(a, n_blocks, s)
    UD2_APL_SWITCH(apl, CALL)
#undef CALL
}
hipError_t launch_lmc_propose(LmcArgs a, int apl, int n_blocks, hipStream_t s) {
#define CALL(N) return launch_lmc_t<N, 1>(a, n_blocks, s)
    UD2_APL_SWITCH(apl, CALL)
#undef CALL
}
hipError_t launch_lmc_step(LmcArgs a, int apl, int n_blocks, hipStream_t s) {
#define CALL(N) return launch_lmc_t<N, 2>(a, n_blocks, s)
    UD2_APL_SWITCH(apl, CALL)
#undef CALL
}

template <int APL>
static hipError_t launch_feat_t(const FeatArgs& a, hipStream_t s) {
    const size_t lds = (size_t)a.lds_floats_per_wave * kWavesPerBlock * sizeof(float);
    const int nb = (a.n_sel + kWavesPerBlock - 1) / kWavesPerBlock;
    hipLaunchKernelGGL(features_kernel<APL>, dim3(nb), dim3(kWave * kWavesPerBlock), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_features(FeatArgs a, int apl, hipStream_t s) {
    if (a.n_sel <= 0) return hipSuccess;
#define CALL(N) return launch_feat_t<N>(a, s)
    UD2_APL_SWITCH(apl, CALL)
#undef CALL
}

template <int APL>
static int lmc_occ_t(int kernel, int lds_bytes) {
    int nb = 0;
    hipError_t e;
    if (kernel == 0)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, lmc_init_kernel<APL>, kWave * kWavesPerBlock, lds_bytes);
    else if (kernel == 1)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, lmc_propose_kernel<APL>, kWave * kWavesPerBlock,
                                                         lds_bytes);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, lmc_step_kernel<APL>, kWave * kWavesPerBlock, lds_bytes);
    if (e != hipSuccess) return 1;
    return nb > 0 ? nb : 1;
}

int lmc_blocks_per_cu(int kernel, int apl, int lds_bytes) {
    switch (apl) {
        case 1: return lmc_occ_t<1>(kernel, lds_bytes);
        case 2: return lmc_occ_t<2>(kernel, lds_bytes);
        case 3: return lmc_occ_t<3>(kernel, lds_bytes);
        default: return lmc_occ_t<4>(kernel, lds_bytes);
    }
}

}  // namespace ud2
