// CPU oracle of the learned MC proposal network (TEST INFRASTRUCTURE ONLY; see
// ud2_oracle.hpp for the rules that apply to every oracle file).
//
// The learned proposal has no reference counterpart: the reference's MC
// (E/src/cuda/mc.cu:146-398) tries one random mutation per step. This file
// restates this project's own definition of the network, so that the device
// kernels can be checked against it:
//
//  * features(): 80 inputs of one candidate pose. For every heavy atom a with
//    feature class k = kFeatClass[type] and its own map value e (trilinear,
//    same clamping as grid_eval) with gradient g, the statistics
//      psi = [1, min(e,0), clamp(e,0,1), log1p(max(e,0)), e < -0.3,
//             min(box overshoot, 4), min(|g|,8)*(e<=0), min(|g|,8)*(e>0)]
//    are summed into features[8k .. 8k+7]; features[64..79] are move globals
//    (kind one-hot, magnitude, ligand size, accepted energy, moved fraction,
//    the curled inter sum) and zeros.
//  * mlp_forward(): y = w3 . relu(W2^T a1 + b2) + b3 with a1 = bf16(relu(W1^T xn + b1)),
//    xn = bf16((x - mu) * inv_sigma); W1, W2, w3 rounded to bf16, biases FP32.
//    The device accumulates in FP32 on MFMA; this restatement accumulates in
//    double and rounds where the device rounds.
//  * mlp_backward(): exact FP64 gradient of the loss w.r.t. every trainable
//    parameter, straight through the bf16 roundings.
//  * group_loss(): rows come in groups of 16 candidates of one MC state. With
//    labels clipped at (group minimum + 6) and r = y_hat - y_clipped, the loss
//    is the mean over rows of (r - mean_group(r))^2: only the order inside a
//    group matters to the MC step, so per-group offsets are not penalised.
//  * adam_step(): Adam (Kingma & Ba, ICLR'15) with bias correction.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "ud2_oracle.hpp"

namespace ud2o {

constexpr int kFeat = 80;        // network inputs
constexpr int kPooled = 64;      // 8 feature classes x 8 statistics
constexpr int kHidden = 256;
constexpr int kGroup = 16;       // candidates per MC state
constexpr float kLabelClip = 6.0f;
// trainable parameters in canonical order: W1[80][256], b1[256], W2[256][256], b2[256], w3[256], b3
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kFeat * kHidden;
constexpr int kOffW2 = kOffB1 + kHidden;
constexpr int kOffB2 = kOffW2 + kHidden * kHidden;
constexpr int kOffW3 = kOffB2 + kHidden;
constexpr int kOffB3 = kOffW3 + kHidden;
constexpr int kNParam = kOffB3 + 1;
// followed by the fixed input normalisation: mu[80], inv_sigma[80]
constexpr int kNNorm = 2 * kFeat;

// Feature class of each XS type (E/src/constants/constants.h:48-68); -1 for hydrogen.
static const int kFeatClass[XS_NTYPES] = {
    -1,  // H
    1,   // B
    0,   // C_H
    1,   // C_P
    2,   // N_P
    3,   // N_D
    4,   // N_A
    3,   // N_DA
    5,   // O_P
    6,   // O_D
    5,   // O_A
    6,   // O_DA
    0,   // F_H
    1,   // Si
    7,   // P_P
    7,   // S_P
    0,   // Cl_H
    0,   // Br_H
    0,   // I_H
    0,   // At
    7,   // Met_D
};

// Round-to-nearest-even to bfloat16, returned as the float it represents.
inline float bf16_round(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) {  // inf / nan: keep, quiet the nan
        if (u & 0x007FFFFFu) u |= 0x00400000u;
        u &= 0xFFFF0000u;
    } else {
        u += 0x7FFFu + ((u >> 16) & 1u);
        u &= 0xFFFF0000u;
    }
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}
inline uint16_t bf16_bits(float f) {
    float r = bf16_round(f);
    uint32_t u;
    std::memcpy(&u, &r, 4);
    return uint16_t(u >> 16);
}
inline float bf16_value(uint16_t b) {
    uint32_t u = uint32_t(b) << 16;
    float r;
    std::memcpy(&r, &u, 4);
    return r;
}

// Move descriptors that enter the global features.
struct MoveInfo {
    int kind = -1;          // -1 none, 0 translation, 1 rotation, 2 torsion
    float magnitude = 0;    // |d| / 2 A, rotation angle / pi, |torsion change| / pi
    float e_acc = 0;        // energy of the accepted MC state
    float moved = 0;        // fraction of heavy atoms the move displaces
};

// Per-atom statistics and pooled / global features of one pose (see the header).
template <class T>
void features(const Ligand& L, const T* x, const Scene& S, const MoveInfo& mv, T out[kFeat]) {
    for (int i = 0; i < kFeat; i++) out[i] = 0;
    const Grid& G = *S.grid;
    T curl_sum = 0;
    int nheavy = 0;
    for (int a = 0; a < L.natom; a++) {
        const int t = L.types[a];
        if (t == XS_H) continue;
        nheavy++;
        const int k = kFeatClass[t];
        const int slot = G.slot_of_type[t];
        T g[3] = {0, 0, 0};
        T e = slot >= 0 ? grid_eval(G, slot, x + 3 * a, g) : T(0);
        T over = 0;
        for (int d = 0; d < 3; d++) {
            if (x[3 * a + d] < T(S.box.lo[d])) over += T(S.box.lo[d]) - x[3 * a + d];
            if (x[3 * a + d] > T(S.box.hi[d])) over += x[3 * a + d] - T(S.box.hi[d]);
        }
        const T gn = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        const T gc = gn < T(8) ? gn : T(8);
        const T ep = e > 0 ? e : T(0);
        T* f = out + 8 * k;
        f[0] += 1;
        f[1] += e < 0 ? e : T(0);
        f[2] += ep < T(1) ? ep : T(1);
        f[3] += std::log1p(ep < T(1e30) ? ep : T(1e30));
        f[4] += e < T(-0.3) ? T(1) : T(0);
        f[5] += over < T(4) ? over : T(4);
        f[6] += e <= 0 ? gc : T(0);
        f[7] += e > 0 ? gc : T(0);
        curl_sum += e > 0 ? e / (T(1) + e) : e;
    }
    T* gl = out + kPooled;
    gl[0] = mv.kind == 0 ? T(1) : T(0);
    gl[1] = mv.kind == 1 ? T(1) : T(0);
    gl[2] = mv.kind == 2 ? T(1) : T(0);
    gl[3] = T(mv.magnitude);
    gl[4] = T(nheavy) / T(64);
    gl[5] = T(L.ntors) / T(16);
    T ea = T(mv.e_acc);
    ea = ea < T(-20) ? T(-20) : (ea > T(20) ? T(20) : ea);
    gl[6] = ea / T(10);
    gl[7] = T(mv.moved);
    T cs = curl_sum < T(-20) ? T(-20) : (curl_sum > T(20) ? T(20) : curl_sum);
    gl[8] = cs / T(10);
}

// Forward pass of nrow rows (x given as bf16 bits, [nrow][80]). round = false
// skips every bf16 rounding (an exact FP64 network, used to pin the backward
// pass by finite differences). Optionally returns the activations for backward.
struct MlpActs {
    std::vector<double> xn, z1, a1, z2, a2;  // per row
};

inline double mlp_forward_row(const float* P, const float* norm, const uint16_t* xrow, bool round,
                              MlpActs* acts = nullptr, const double* xin_exact = nullptr) {
    const float* mu = norm;
    const float* inv = norm + kFeat;
    auto rw = [&](float w) -> double { return round ? double(bf16_round(w)) : double(w); };
    double xn[kFeat];
    for (int f = 0; f < kFeat; f++) {
        if (xin_exact) {
            xn[f] = (xin_exact[f] - double(mu[f])) * double(inv[f]);
        } else {
            const float xv = bf16_value(xrow[f]);
            const float v = (xv - mu[f]) * inv[f];
            xn[f] = round ? double(bf16_round(v)) : double(v);
        }
    }
    double z1[kHidden], a1[kHidden], z2[kHidden], a2[kHidden];
    for (int h = 0; h < kHidden; h++) {
        double s = 0;
        for (int f = 0; f < kFeat; f++) s += rw(P[kOffW1 + f * kHidden + h]) * xn[f];
        z1[h] = s + double(P[kOffB1 + h]);
        const double r = z1[h] > 0 ? z1[h] : 0.0;
        a1[h] = round ? double(bf16_round(float(r))) : r;
    }
    double y = double(P[kOffB3]);
    for (int o = 0; o < kHidden; o++) {
        double s = 0;
        for (int h = 0; h < kHidden; h++) s += rw(P[kOffW2 + h * kHidden + o]) * a1[h];
        z2[o] = s + double(P[kOffB2 + o]);
        a2[o] = z2[o] > 0 ? z2[o] : 0.0;
        y += rw(P[kOffW3 + o]) * a2[o];
    }
    if (acts) {
        acts->xn.insert(acts->xn.end(), xn, xn + kFeat);
        acts->z1.insert(acts->z1.end(), z1, z1 + kHidden);
        acts->a1.insert(acts->a1.end(), a1, a1 + kHidden);
        acts->z2.insert(acts->z2.end(), z2, z2 + kHidden);
        acts->a2.insert(acts->a2.end(), a2, a2 + kHidden);
    }
    return y;
}

// Gradient of sum_r dy[r] * y_hat[r] w.r.t. the trainable parameters (canonical
// order), accumulated in double into grad[kNParam].
inline void mlp_backward(const float* P, const float* norm, int nrow, const uint16_t* x, const double* dy,
                         bool round, double* grad) {
    auto rw = [&](float w) -> double { return round ? double(bf16_round(w)) : double(w); };
    for (int i = 0; i < kNParam; i++) grad[i] = 0;
    std::vector<double> dz2(kHidden), dz1(kHidden);
    for (int r = 0; r < nrow; r++) {
        MlpActs A;
        mlp_forward_row(P, norm, x + size_t(r) * kFeat, round, &A);
        const double g = dy[r];
        grad[kOffB3] += g;
        for (int o = 0; o < kHidden; o++) {
            grad[kOffW3 + o] += g * A.a2[o];
            dz2[o] = A.z2[o] > 0 ? g * rw(P[kOffW3 + o]) : 0.0;
            grad[kOffB2 + o] += dz2[o];
        }
        for (int h = 0; h < kHidden; h++) {
            double s = 0;
            for (int o = 0; o < kHidden; o++) {
                s += rw(P[kOffW2 + h * kHidden + o]) * dz2[o];
                grad[kOffW2 + h * kHidden + o] += A.a1[h] * dz2[o];
            }
            dz1[h] = A.z1[h] > 0 ? s : 0.0;
            grad[kOffB1 + h] += dz1[h];
        }
        for (int f = 0; f < kFeat; f++)
            for (int h = 0; h < kHidden; h++) grad[kOffW1 + f * kHidden + h] += A.xn[f] * dz1[h];
    }
}

// Group-centred loss over ngroup groups of kGroup rows; returns the loss and
// writes d loss / d y_hat.
inline double group_loss(int ngroup, const double* yhat, const float* label, double* dyhat) {
    const int n = ngroup * kGroup;
    double loss = 0;
    for (int g = 0; g < ngroup; g++) {
        const double* yh = yhat + g * kGroup;
        const float* y = label + g * kGroup;
        double ymin = INFINITY;
        for (int i = 0; i < kGroup; i++)
            if (std::isfinite(y[i]) && double(y[i]) < ymin) ymin = double(y[i]);
        double r[kGroup], rbar = 0;
        for (int i = 0; i < kGroup; i++) {
            double yc = std::isfinite(y[i]) ? double(y[i]) : INFINITY;
            if (yc > ymin + kLabelClip) yc = ymin + kLabelClip;
            r[i] = yh[i] - yc;
            rbar += r[i];
        }
        rbar /= kGroup;
        for (int i = 0; i < kGroup; i++) {
            const double c = std::isfinite(ymin) ? r[i] - rbar : 0.0;
            loss += c * c;
            dyhat[g * kGroup + i] = 2.0 * c / n;
        }
    }
    return loss / n;
}

// One Adam step on every trainable parameter, in float as the device does.
inline void adam_step(float* P, float* m, float* v, const float* grad, int n, float lr, float beta1, float beta2,
                      float eps, int step) {
    const float c1 = 1.0f - std::pow(beta1, float(step));
    const float c2 = 1.0f - std::pow(beta2, float(step));
    for (int i = 0; i < n; i++) {
        const float g = grad[i];
        m[i] = beta1 * m[i] + (1.0f - beta1) * g;
        v[i] = beta2 * v[i] + (1.0f - beta2) * g * g;
        const float mh = m[i] / c1, vh = v[i] / c2;
        P[i] -= lr * mh / (std::sqrt(vh) + eps);
    }
}

}  // namespace ud2o
