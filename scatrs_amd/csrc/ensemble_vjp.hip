// Reverse mode of the fixed-grid ensemble path (scatrs_ensemble_fixed_vjp): the vector-Jacobian
// product of the scalar, Greek and phase outputs with respect to each ensemble's distribution
// parameters, layer indices and radius fractions.
//
// The loss of one ensemble is L = sum_k Sbar_k S_k + sum_{k,j} Tbar_kj T_kj over the sums the
// forward kernels form, S = sum_i W_i h_i (h = [1, C_ext, C_sca, C_sca g, C_abs, C_back, r^2,
// r^3]) and T = sum_i W_i (lambda^2 / pi) F_i (the table factor pi r^2 4 / x^2 is
// lambda^2 / pi). Every cross section is lambda^2 times a function of a_n, b_n alone (C_abs of
// the absorption sum), so a node's share phi_i = W_i H_i depends on its radius only through
// the coefficients, r^2, r^3 and the weight. Per chunk:
//   forward_nodes, forward_angular - the forward node solves and angular products
//                      (coefficients staged in FP64, per-node backscatter sums kept; with the
//                      fp16-split kernel the products read an fp32 copy, stage_fp32);
//   ens_greek_vjp_mfma - the transposed Greek projection of the Greek cotangents;
//   ens_finish_vjp   - per ensemble, the reverse of ens_finish: Sbar (8) and Tbar (4 x angles);
//   ens_angular_vjp  - per (ensemble, 16 nodes): S1, S2 at every cosine again on the FP64
//                      matrix cores, the cotangents of S1, S2 from Tbar, and the transposed
//                      products back onto the orders: the table part of each node's a_n, b_n
//                      cotangents, and the node's sum Tbar . F;
//   ens_nodes_vjp_ranked - per node (the forward chunk's ranking and launches), a_n, b_n and
//                      their derivatives recomputed (not stored): the
//                      cotangents completed with the scalar outputs' part, contracted with the
//                      coefficient derivatives (and, for C_abs, the absorption sum's closed-form
//                      Jacobian);
//   ens_node_grads   - per node, those contractions with the node's weight and radius
//                      derivatives: its share of the gradient;
//   ens_grad_reduce  - per ensemble, the nodes' gradients summed in node order.
#include "adjoint_device.h"
#include "ensemble_kernels.h"

namespace scatrs {

// ---- the reverse of ens_finish ----

// Sum of one value per thread over a block of 256, in a fixed tree order.
static __device__ double block_sum(double v, double *red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) red[t] += red[t + off];
        __syncthreads();
    }
    const double s = red[0];
    __syncthreads();
    return s;
}

// From the output cotangents of one ensemble (bar_scalars (11), bar_greek (6 x (order + 1)),
// bar_phase (6 x n_angles); each may be null) those of the sums, sbar[eb * 8 + k], and of the
// summed tables, tbar[(eb * 4 + k) * n_angles + j]. Follows ens_finish step by step in reverse
// (ensemble/engine.rs:1782-1836,2396-2497 for the forward steps). With n_out > 0 the phase
// cotangents are at the caller's n_mu cosines: tab_u holds the tables at the n_out cosines
// -|mu|, +|mu| (ens_phase_mu's layout) and position q gathers the cosines
// q_list[q_ptr[q] .. q_ptr[q + 1]); their cotangents go to tbar_u in that layout.
static __global__ __launch_bounds__(256) void ens_finish_vjp(
    int64_t e0, int64_t n_radius, int n_angles, const double *aw, const double *vals,
    const double *tab, const double *dtab, int64_t order, int normalized,
    const double *bar_scalars, const double *bar_greek, const double *bar_phase, double *sbar,
    double *tbar, int n_out, int n_mu, const double *tab_u, const int *q_ptr, const int *q_list,
    double *tbar_u) {
    extern __shared__ double sh[];
    double *S = sh;      // 8 sums
    double *F = sh + 8;  // 4 * n_angles, divided by Z
    __shared__ double red[256];
    const int64_t e = e0 + blockIdx.x;
    const int t = threadIdx.x;
    const double *v = vals + blockIdx.x * n_radius * kNodeVals;
    if (t < kNodeVals) {
        double s = 0.0;
        for (int64_t i = 0; i < n_radius; ++i) s += v[i * kNodeVals + t];
        S[t] = s;
    }
    __syncthreads();
    const double z = S[0];
    const double *te = tab + (int64_t)blockIdx.x * 4 * n_angles;
    for (int j = t; j < 4 * n_angles; j += blockDim.x) F[j] = te[j] / z;
    __syncthreads();
    const double ext = S[1] / z, sca = S[2] / z, scag = S[3] / z, back = S[5] / z, r2 = S[6] / z,
                 r3 = S[7] / z;
    const double *F11 = F, *F12 = F + n_angles, *F33 = F + 2 * n_angles, *F34 = F + 3 * n_angles;
    const int64_t len = order + 1;
    const bool greek = bar_greek && dtab;
    double *tb = tbar + (int64_t)blockIdx.x * 4 * n_angles;
    // With Greek cotangents, tb holds their transposed projection B (ens_greek_vjp_mfma), the
    // cotangent of the weighted tables w F through gamma_kl = c_l sum_j w_j (F . d)_kl. When
    // normalized, G = gamma / alpha1_0 with alpha1_0 = gamma_00: the outputs' cotangents act
    // through gamma / alpha1_0 (B / alpha1_0) and through alpha1_0, whose cotangent is
    // -sum_{kl != 00} gbar_kl gamma_kl / alpha1_0^2 = -(B . F) / alpha1_0^2 (B leaves gbar_00 out).
    double a0 = 1.0, a0bar = 0.0;
    if (greek && normalized) {
        double a = 0.0, part = 0.0;
        for (int j = t; j < n_angles; j += blockDim.x) {
            a += (aw[j] * F11[j]) * dtab[(int64_t)j * len];
            part += tb[j] * F11[j] + tb[n_angles + j] * F12[j] + tb[2 * n_angles + j] * F33[j] +
                    tb[3 * n_angles + j] * F34[j];
        }
        a0 = 0.5 * block_sum(a, red);
        a0bar = -block_sum(part, red) / (a0 * a0);
    }
    // Phase: ph = src / sca_n with sca_n = <C_sca> when normalized.
    const double sca_n = normalized ? sca : 1.0;
    const double *pb = bar_phase && n_out == 0 ? bar_phase + e * 6 * (int64_t)n_angles : nullptr;
    double fbf = 0.0, ph_sca = 0.0;
    if (n_out > 0 && bar_phase) {
        // The caller's cosines: position q of the symmetric layout, its tables F_u = T_u / Z.
        const double *pu = bar_phase + e * 6 * (int64_t)n_mu;
        const double *tu = tab_u + (int64_t)blockIdx.x * 4 * n_out;
        double *tbu = tbar_u + (int64_t)blockIdx.x * 4 * n_out;
        for (int q = t; q < n_out; q += blockDim.x) {
            const double f11 = tu[q] / z, f12 = tu[n_out + q] / z, f33 = tu[2 * n_out + q] / z,
                         f34 = tu[3 * n_out + q] / z;
            double b11 = 0.0, b12 = 0.0, b33 = 0.0, b34 = 0.0;
            for (int k = q_ptr[q]; k < q_ptr[q + 1]; ++k) {
                const int j = q_list[k];
                const double p11 = pu[0 * n_mu + j] + pu[2 * n_mu + j];
                const double p12 = pu[1 * n_mu + j];
                const double p33 = pu[3 * n_mu + j] + pu[5 * n_mu + j];
                const double p34 = pu[4 * n_mu + j];
                b11 += p11 / sca_n;
                b12 += p12 / sca_n;
                b33 += p33 / sca_n;
                b34 += p34 / sca_n;
                if (normalized)
                    ph_sca -= (p11 * f11 + p12 * f12 + p33 * f33 + p34 * f34) / (sca * sca);
            }
            fbf += b11 * f11 + b12 * f12 + b33 * f33 + b34 * f34;
            tbu[q] = b11 / z;
            tbu[n_out + q] = b12 / z;
            tbu[2 * n_out + q] = b33 / z;
            tbu[3 * n_out + q] = b34 / z;
        }
    }
    for (int j = t; j < n_angles; j += blockDim.x) {
        double b11 = 0.0, b12 = 0.0, b33 = 0.0, b34 = 0.0;
        if (greek) {
            const double ia0 = normalized ? 1.0 / a0 : 1.0;
            b11 = tb[j] * ia0;
            b12 = tb[n_angles + j] * ia0;
            b33 = tb[2 * n_angles + j] * ia0;
            b34 = tb[3 * n_angles + j] * ia0;
            // alpha1_0's cotangent through gamma_00 = (1/2) sum_j w_j F11_j d00_0(mu_j).
            if (normalized) b11 += aw[j] * 0.5 * a0bar * dtab[(int64_t)j * len];
        }
        if (pb) {
            const double p11 = pb[0 * n_angles + j] + pb[2 * n_angles + j];
            const double p12 = pb[1 * n_angles + j];
            const double p33 = pb[3 * n_angles + j] + pb[5 * n_angles + j];
            const double p34 = pb[4 * n_angles + j];
            b11 += p11 / sca_n;
            b12 += p12 / sca_n;
            b33 += p33 / sca_n;
            b34 += p34 / sca_n;
            if (normalized)
                ph_sca -= (p11 * F11[j] + p12 * F12[j] + p33 * F33[j] + p34 * F34[j]) / (sca * sca);
        }
        fbf += b11 * F11[j] + b12 * F12[j] + b33 * F33[j] + b34 * F34[j];
        tb[0 * n_angles + j] = b11 / z;
        tb[1 * n_angles + j] = b12 / z;
        tb[2 * n_angles + j] = b33 / z;
        tb[3 * n_angles + j] = b34 / z;
    }
    fbf = block_sum(fbf, red);
    ph_sca = block_sum(ph_sca, red);
    if (t == 0) {
        double ob[11] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (bar_scalars)
            for (int k = 0; k < 11; ++k) ob[k] = bar_scalars[e * 11 + k];
        // o = [ext, sca, abs, back, sca/ext, scag/sca, 4 pi ext/back, 0, r3/r2, pi r2, 4pi/3 r3]
        const double q_ext = ob[0] - ob[4] * sca / (ext * ext) + ob[6] * 4.0 * kPi / back;
        const double q_sca = ob[1] + ob[4] / ext - ob[5] * scag / (sca * sca) + ph_sca;
        const double q_scag = ob[5] / sca;
        const double q_abs = ob[2];
        const double q_back = ob[3] - ob[6] * 4.0 * kPi * ext / (back * back);
        const double q_r2 = -ob[8] * r3 / (r2 * r2) + ob[9] * kPi;
        const double q_r3 = ob[8] / r2 + ob[10] * 4.0 * kPi / 3.0;
        const double q[7] = {q_ext, q_sca, q_scag, q_abs, q_back, q_r2, q_r3};
        double zbar = -fbf;
        double *sb = sbar + blockIdx.x * 8;
        for (int k = 0; k < 7; ++k) {
            zbar -= q[k] * (S[k + 1] / z);
            sb[k + 1] = q[k] / z;
        }
        sb[0] = zbar / z;
    }
}

// The transpose of ens_greek_mfma, for the Greek cotangents gbar (count, 6, len) of a chunk:
// per ensemble and angle, out = [w (Y0 + Y2 + Y3), w Y4, w (Y1 + Y2 - Y3), w Y5] in the table
// components' order (F11, F12, F33, F34), Y_k[e][j] = sum_l U_k[e][l] D_k[l][j] with
// U = c_l (gbar_0, gbar_3, (gbar_1 + gbar_2)/2, (gbar_1 - gbar_2)/2, gbar_4, gbar_5),
// c_l = (2l + 1)/2 and D = (d00, d00, d22, d2-2, d02, d02), as GEMMs in tiles of 16 ensembles x
// 16 angles with v_mfma_f64_16x16x4f64 (operand and result lanes as in ens_greek_mfma). A
// workgroup is 16 ensembles and 64 angles, a wavefront 16 angles; the U tile of 64 orders is
// staged in LDS, the table read order-major (dtab_t[(p * len + l) * n_angles + j]), so that a
// wavefront's loads run along the angles. With normalization gbar_0 at l = 0 is left out
// (alpha1_0 is 1 exactly); ens_finish_vjp adds the cotangent of alpha1_0 in its place.
constexpr int kGreekOrders = 64;  // orders of the U tile staged per step

static __global__ __launch_bounds__(256) void ens_greek_vjp_mfma(
    int64_t e0, int64_t ne, int n_angles, const double *aw, const double *bar_greek,
    const double *dtab_t, int64_t len, int normalized, double *out) {
    __shared__ double lds_u[6][16][kGreekOrders + 1];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int r = lane & 15, kq = lane >> 4;
    const int64_t et = (int64_t)blockIdx.x * 16;                // the tile's first ensemble
    const int jb = (int)blockIdx.y * 64 + wave * 16 + r;         // this lane's B column (angle)
    const double *dcol = dtab_t + (jb < n_angles ? jb : 0);
    const int64_t plane = len * n_angles;
    mfma_d4 acc[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[k] = mfma_d4{0, 0, 0, 0};
    for (int64_t l0 = 0; l0 < len; l0 += kGreekOrders) {
        __syncthreads();
        // Thread t stages (ensemble t / 16, orders t % 16 + 16 q) of the tile.
#pragma unroll
        for (int q = 0; q < kGreekOrders / 16; ++q) {
            const int er = t / 16, la = t % 16 + 16 * q;
            const int64_t e = et + er, l = l0 + la;
            double g[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            if (e < ne && l < len) {  // rows beyond the chunk, orders beyond len: 0
                const double *gb = bar_greek + (e0 + e) * 6 * len + l;
                const double c = (double)(2 * l + 1) * 0.5;
#pragma unroll
                for (int k = 0; k < 6; ++k) g[k] = c * gb[k * len];
                if (normalized && l == 0) g[0] = 0.0;
            }
            lds_u[0][er][la] = g[0];
            lds_u[1][er][la] = g[3];
            lds_u[2][er][la] = 0.5 * (g[1] + g[2]);
            lds_u[3][er][la] = 0.5 * (g[1] - g[2]);
            lds_u[4][er][la] = g[4];
            lds_u[5][er][la] = g[5];
        }
        __syncthreads();
#pragma unroll 4
        for (int ll = 0; ll < kGreekOrders; ll += 4) {
            const int la = ll + kq;
            const int64_t l = l0 + la;
            // Orders beyond len meet zero U values; read a valid row of the table for them.
            const double *dl = dcol + (l < len ? l : 0) * n_angles;
            const double d00 = dl[0], d22 = dl[plane], d2m2 = dl[2 * plane], d02 = dl[3 * plane];
            acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(lds_u[0][r][la], d00, acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(lds_u[1][r][la], d00, acc[1], 0, 0, 0);
            acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(lds_u[2][r][la], d22, acc[2], 0, 0, 0);
            acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(lds_u[3][r][la], d2m2, acc[3], 0, 0, 0);
            acc[4] = __builtin_amdgcn_mfma_f64_16x16x4f64(lds_u[4][r][la], d02, acc[4], 0, 0, 0);
            acc[5] = __builtin_amdgcn_mfma_f64_16x16x4f64(lds_u[5][r][la], d02, acc[5], 0, 0, 0);
        }
    }
    if (jb >= n_angles) return;
    const double w = aw[jb];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int64_t e = et + kq + 4 * v;
        if (e >= ne) continue;
        double *o = out + e * 4 * (int64_t)n_angles + jb;
        o[0] = w * (acc[0][v] + acc[2][v] + acc[3][v]);
        o[n_angles] = w * acc[4][v];
        o[2 * n_angles] = w * (acc[1][v] + acc[2][v] - acc[3][v]);
        o[3 * n_angles] = w * acc[5][v];
    }
}

// ---- the angular products in reverse ----

constexpr int kVjpTiles = 2;                // MFMA tiles of 4 nodes per group
constexpr int kVjpNodes = 4 * kVjpTiles;    // nodes per group
constexpr int kVjpCols = 16 * kVjpTiles;    // matrix-core columns of a group: 4 per node
constexpr int kVjpOrders = 64;  // coefficient chunk of the forward products
constexpr int kVjpAngles = 64;  // panel chunk of the transposed products
constexpr int kVjpOrderRows = 128;

// Persistent: each workgroup takes (ensemble, group of kVjpNodes nodes) items from `counter`
// (groups of 8 nodes keep a wavefront's accumulators within 128 registers: 4 wavefronts per SIMD).
//  1. S1, S2 at every non-negative cosine and its mirror (as ens_angular_mfma, but per node).
//  2. With Tbar at +-mu, g1 = dL/dS1 = (b11 - b12) S1 + (b33 - i b34) S2 and
//     g2 = dL/dS2 = (b11 + b12) S2 + (b33 + i b34) S1 (L = Tbar . F as a real function of the
//     real and imaginary parts); the node's Tbar . F is summed on the way.
//  3. S1(+-mu) = A +- B, S2(+-mu) = C +- D with A, C = G1 u, G1 v and B, D = G2 v, G2 u give
//     ubar = G1^T (g1+ + g1-) + G2^T (g2+ - g2-), vbar = G2^T (g1+ - g1-) + G1^T (g2+ + g2-),
//     the cotangents (dL = Re sum conj(ubar) du) of u, v = a_n, b_n (odd n) or b_n, a_n (even n).
//     The panels P = [g1+ + g1-, g2+ + g2-] and M = [g2+ - g2-, g1+ - g1-] (re, im per node)
//     are written to the workgroup's slab, and two transposed products per 128 orders form
//     G1^T P + G2^T M: columns 4q + c hold node q's (ubar.re, ubar.im, vbar.re, vbar.im)[c].
//     A wavefront whose 16 orders lie beyond the group's largest N skips its products.
// With accumulate set the cotangents and Tbar . F add to what abar and ht hold (a second pass
// at the caller's phase cosines after the rule's).
static __global__ __launch_bounds__(512, 4) void ens_angular_vjp(
    int64_t e0, int64_t ne, int64_t n_radius, int n_angles, int h, const int *node_n,
    const int64_t *node_off, const int64_t *ens_base, const cx *staging, const double2 *table,
    const double *tbar, unsigned long long *counter, double *slabs, cx *abar, double *ht,
    int accumulate) {
    // The chunk's coefficients as the A operand: [tile][row 4 c + q][order], u, v by parity.
    __shared__ double lds_v[kVjpTiles][16][kVjpOrders + 1];
    __shared__ double lds_g[2][kVjpAngles][kVjpCols];
    __shared__ double lds_ht[8][kVjpNodes];
    __shared__ unsigned long long item_sh;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int row = lane & 15, kq = lane >> 4;
    const int comp = lane & 3, qin = (lane & 15) >> 2;  // transposed products' D column
    const int sq = t / kVjpOrders, sk = t % kVjpOrders;  // this thread's staging pair
    double *slab = slabs + (size_t)blockIdx.x * 2 * kVjpCols * h;
    const int64_t groups = (n_radius + kVjpNodes - 1) / kVjpNodes;
    const int64_t items = ne * groups;
    while (true) {
        __syncthreads();  // the previous item is done with the LDS and the slab
        if (t == 0) item_sh = atomicAdd(counter, 1ull);
        __syncthreads();
        const int64_t item = (int64_t)item_sh;
        if (item >= items) break;
        const int64_t eb = item / groups, g0 = (item % groups) * kVjpNodes;
        const int64_t e = e0 + eb;
        const int64_t base = ens_base[e];
        const int *nn = node_n + e * n_radius;
        const int64_t *noff = node_off + e * n_radius;
        const double *tb = tbar + eb * 4 * (int64_t)n_angles;
        int nmax = 0;
        for (int q = 0; q < kVjpNodes; ++q) {
            const int nq = g0 + q < n_radius ? nn[g0 + q] : 0;
            nmax = nq > nmax ? nq : nmax;
        }
        double hts[kVjpTiles] = {};  // Tbar . F of node 4 tt + kq over this lane's cosines
        // 1-2: forward products per slice of 128 cosines, cotangent panels to the slab.
        for (int slice = 0; slice * kSliceAngles < h; ++slice) {
            const int jj = slice * kSliceAngles + wave * 16 + row;
            mfma_d4 acc1[kVjpTiles], acc2[kVjpTiles];
#pragma unroll
            for (int tt = 0; tt < kVjpTiles; ++tt) {
                acc1[tt] = mfma_d4{0, 0, 0, 0};
                acc2[tt] = mfma_d4{0, 0, 0, 0};
            }
            // Pipelined: the next chunk's coefficient pair of this thread (node sq, order sk)
            // loads while a chunk's products run, and the table two steps ahead of its use
            // (from a clamped address, masked after arrival); the products and their order are
            // those of the unpipelined loop.
            cx pa = real(0.0), pb = real(0.0);
            auto load_pair = [&](int kc) __attribute__((always_inline)) {
                const int n = kc + sk;
                const int nq = g0 + sq < n_radius ? nn[g0 + sq] : 0;
                pa = real(0.0);
                pb = real(0.0);
                if (n < nq) {
                    const cx *c = staging + 2 * (base + noff[g0 + sq]);
                    pa = c[2 * n];
                    pb = c[2 * n + 1];
                }
            };
            auto tab = [&](int kc, int st) __attribute__((always_inline)) {
                const int n = kc + 4 * st + kq + 1;
                const bool ok = jj < h && n <= nmax;
                const double2 v = table[(size_t)(ok ? n - 1 : 0) * h + (jj < h ? jj : 0)];
                return ok ? v : make_double2(0.0, 0.0);
            };
            if (nmax > 0) load_pair(0);
            for (int kc = 0; kc < nmax; kc += kVjpOrders) {
                __syncthreads();
                {
                    const int n = kc + sk;
                    const bool odd = (n + 1) & 1;
                    const cx u = odd ? pa : pb, v = odd ? pb : pa;
                    double *col = &lds_v[sq / 4][sq % 4][sk];
                    col[0] = u.re;
                    col[4 * (kVjpOrders + 1)] = u.im;
                    col[8 * (kVjpOrders + 1)] = v.re;
                    col[12 * (kVjpOrders + 1)] = v.im;
                }
                __syncthreads();
                if (kc + kVjpOrders < nmax) load_pair(kc + kVjpOrders);
                const int kend = nmax - kc < kVjpOrders ? nmax - kc : kVjpOrders;
                double2 r0 = tab(kc, 0), r1 = tab(kc, 1);
                for (int k4 = 0; k4 < kend; k4 += 4) {
                    const int kk = k4 + kq;
                    const int n = kc + kk + 1;
                    const double2 pt = r0;
                    r0 = r1;
                    r1 = tab(kc, k4 / 4 + 2);
                    const bool odd = n & 1;
                    const double g1 = odd ? pt.x : pt.y;
                    const double g2 = odd ? pt.y : pt.x;
#pragma unroll
                    for (int tt = 0; tt < kVjpTiles; ++tt) {
                        // A row 4 c + q: component c (u.re, u.im, v.re, v.im) of node 4 tt + q.
                        const double val = lds_v[tt][row][kk];
                        acc1[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(val, g1, acc1[tt], 0, 0, 0);
                        acc2[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(val, g2, acc2[tt], 0, 0, 0);
                    }
                }
            }
            // D[4 c + q][j] sits in lane j + 16 q, element c: this lane holds the four components
            // of node 4 tt + kq at its cosine (row).
            if (jj < h) {
                const int jp = n_angles - h + jj, jm = n_angles - 1 - jp;
                double bp[4], bm[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    bp[k] = tb[k * n_angles + jp];
                    bm[k] = jm != jp ? tb[k * n_angles + jm] : 0.0;
                }
#pragma unroll
                for (int tt = 0; tt < kVjpTiles; ++tt) {
                    const cx A = mk(acc1[tt][0], acc1[tt][1]), C = mk(acc1[tt][2], acc1[tt][3]);
                    const cx D = mk(acc2[tt][0], acc2[tt][1]), B = mk(acc2[tt][2], acc2[tt][3]);
                    const cx s1p = A + B, s1m = A - B, s2p = C + D, s2m = C - D;
                    const cx g1p = s1p * (bp[0] - bp[1]) + s2p * bp[2] - times_i(s2p) * bp[3];
                    const cx g2p = s2p * (bp[0] + bp[1]) + s1p * bp[2] + times_i(s1p) * bp[3];
                    const cx g1m = s1m * (bm[0] - bm[1]) + s2m * bm[2] - times_i(s2m) * bm[3];
                    const cx g2m = s2m * (bm[0] + bm[1]) + s1m * bm[2] + times_i(s1m) * bm[3];
                    const cx P1 = g1p + g1m, M1 = g1p - g1m, P2 = g2p + g2m, M2 = g2p - g2m;
                    double *sp = slab + (size_t)jj * kVjpCols + 16 * tt + 4 * kq;
                    double *sm = slab + ((size_t)h + jj) * kVjpCols + 16 * tt + 4 * kq;
                    sp[0] = P1.re;
                    sp[1] = P1.im;
                    sp[2] = P2.re;
                    sp[3] = P2.im;
                    sm[0] = M2.re;
                    sm[1] = M2.im;
                    sm[2] = M1.re;
                    sm[3] = M1.im;
                    // F11, F12, F33, F34 at +mu and -mu against their cotangents
                    const double n1p = norm2(s1p), n2p = norm2(s2p), n1m = norm2(s1m), n2m = norm2(s2m);
                    const cx pp = s2p * conj(s1p), pm = s2m * conj(s1m);
                    hts[tt] += bp[0] * (0.5 * (n1p + n2p)) + bp[1] * (0.5 * (n2p - n1p)) +
                               bp[2] * pp.re + bp[3] * pp.im;
                    hts[tt] += bm[0] * (0.5 * (n1m + n2m)) + bm[1] * (0.5 * (n2m - n1m)) +
                               bm[2] * pm.re + bm[3] * pm.im;
                }
            }
        }
        // The node's Tbar . F: over the 16 cosine lanes of the node, then over the waves.
#pragma unroll
        for (int tt = 0; tt < kVjpTiles; ++tt) {
            double x = hts[tt];
            x += __shfl_xor(x, 1, 64);
            x += __shfl_xor(x, 2, 64);
            x += __shfl_xor(x, 4, 64);
            x += __shfl_xor(x, 8, 64);
            hts[tt] = x;
        }
        if (row == 0)
#pragma unroll
            for (int tt = 0; tt < kVjpTiles; ++tt) lds_ht[wave][4 * tt + kq] = hts[tt];
        __syncthreads();  // the slab's panels and lds_ht are complete
        if (t < kVjpNodes && g0 + t < n_radius) {
            double s = 0.0;
            for (int w = 0; w < 8; ++w) s += lds_ht[w][t];
            double *o = ht + eb * n_radius + g0 + t;
            *o = accumulate ? *o + s : s;
        }
        // 3: the transposed products, 128 orders at a time in ob <= 8 blocks of 16. Wavefront
        // w takes order block w % ob and angle split w / ob, for the splits = 8, 4, 2 or 1
        // (ob = 1, 2, 3-4, 5-8) that the 8 wavefronts hold: each split takes a contiguous run of
        // every panel chunk's angle steps, and the splits' partial sums are added in split
        // order through LDS, so that a small group's orders do not leave wavefronts idle.
        for (int nb = 0; nb < nmax; nb += kVjpOrderRows) {
            const int rem = nmax - nb;
            const int ob = rem >= kVjpOrderRows ? 8 : (rem + 15) / 16;
            const int splits = ob == 1 ? 8 : ob == 2 ? 4 : ob <= 4 ? 2 : 1;
            const int obi = wave % ob, sp = wave / ob;
            const bool active = sp < splits;
            const int sp_angles = kVjpAngles / splits;   // a split's angles of each chunk
            const int n_row = nb + obi * 16 + row + 1;  // this lane's A-operand order
            const bool odd_row = n_row & 1;
            mfma_d4 d1[kVjpTiles], d2[kVjpTiles];
#pragma unroll
            for (int tt = 0; tt < kVjpTiles; ++tt) {
                d1[tt] = mfma_d4{0, 0, 0, 0};
                d2[tt] = mfma_d4{0, 0, 0, 0};
            }
            // Pipelined as the forward products: the next panel chunk (this thread's kPanelVals
            // values) loads while a chunk's products run, the table two steps ahead.
            constexpr int kPanelVals = 2 * kVjpAngles * kVjpCols / 512;
            double pv[kPanelVals];
            auto load_panel = [&](int jc) __attribute__((always_inline)) {
#pragma unroll
                for (int r = 0; r < kPanelVals; ++r) {
                    const int idx = t + 512 * r;
                    const int s = idx / (kVjpAngles * kVjpCols), k = (idx / kVjpCols) % kVjpAngles,
                              col = idx % kVjpCols;
                    const int j = jc + k;
                    pv[r] = j < h ? slab[((size_t)s * h + j) * kVjpCols + col] : 0.0;
                }
            };
            auto tabr = [&](int jc, int st) __attribute__((always_inline)) {
                const int jb = jc + 4 * st + kq;
                const bool ok = n_row <= nmax && jb < h;
                const double2 v = table[(size_t)(ok ? n_row - 1 : 0) * h + (ok ? jb : 0)];
                return ok ? v : make_double2(0.0, 0.0);
            };
            load_panel(0);
            for (int jc = 0; jc < h; jc += kVjpAngles) {
                __syncthreads();
#pragma unroll
                for (int r = 0; r < kPanelVals; ++r) {
                    const int idx = t + 512 * r;
                    const int s = idx / (kVjpAngles * kVjpCols), k = (idx / kVjpCols) % kVjpAngles,
                              col = idx % kVjpCols;
                    lds_g[s][k][col] = pv[r];
                }
                __syncthreads();
                if (jc + kVjpAngles < h) load_panel(jc + kVjpAngles);
                const int cend = h - jc < kVjpAngles ? h - jc : kVjpAngles;
                const int kbeg = sp * sp_angles;
                const int kend = !active ? 0 : (kbeg + sp_angles < cend ? kbeg + sp_angles : cend);
                double2 r0 = make_double2(0.0, 0.0), r1 = r0;
                if (kbeg < kend) {
                    r0 = tabr(jc, kbeg / 4);
                    r1 = tabr(jc, kbeg / 4 + 1);
                }
                for (int k4 = kbeg; k4 < kend; k4 += 4) {
                    const double2 pt = r0;
                    r0 = r1;
                    r1 = tabr(jc, k4 / 4 + 2);
                    const double g1 = odd_row ? pt.x : pt.y;
                    const double g2 = odd_row ? pt.y : pt.x;
#pragma unroll
                    for (int tt = 0; tt < kVjpTiles; ++tt) {
                        const double b1 = lds_g[0][k4 + kq][16 * tt + row];
                        const double b2 = lds_g[1][k4 + kq][16 * tt + row];
                        d1[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(g1, b1, d1[tt], 0, 0, 0);
                        d2[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(g2, b2, d2[tt], 0, 0, 0);
                    }
                }
            }
            double dv[kVjpTiles][4];
#pragma unroll
            for (int tt = 0; tt < kVjpTiles; ++tt)
#pragma unroll
                for (int v = 0; v < 4; ++v) dv[tt][v] = d1[tt][v] + d2[tt][v];
            if (splits > 1) {
                // The splits' partial sums through the panel tile (8 x 8 x 64 doubles).
                double *red = &lds_g[0][0][0];
                __syncthreads();  // every wavefront is done with the last panel chunk
                if (active)
#pragma unroll
                    for (int tt = 0; tt < kVjpTiles; ++tt)
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            red[((wave * kVjpTiles + tt) * 4 + v) * 64 + lane] = dv[tt][v];
                __syncthreads();
                if (active && sp == 0)
#pragma unroll
                    for (int tt = 0; tt < kVjpTiles; ++tt)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            double x = dv[tt][v];
                            for (int q = 1; q < splits; ++q)
                                x += red[(((q * ob + obi) * kVjpTiles + tt) * 4 + v) * 64 + lane];
                            dv[tt][v] = x;
                        }
            }
            if (!active || sp != 0) continue;
#pragma unroll
            for (int tt = 0; tt < kVjpTiles; ++tt) {
                const int64_t node = g0 + 4 * tt + qin;
                if (node >= n_radius) continue;
                const int nq = nn[node];
                cx *cb = abar + 2 * (base + noff[node]);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int n = nb + obi * 16 + kq + 4 * v + 1;
                    if (n > nq) continue;
                    const double val = dv[tt][v];
                    const bool to_a = (comp < 2) == (bool)(n & 1);  // u is a_n at odd n
                    double *dst = reinterpret_cast<double *>(cb + (to_a ? 0 : nq) + n - 1);
                    dst[comp & 1] = accumulate ? dst[comp & 1] + val : val;
                }
            }
        }
    }
}

// ---- the node solves in reverse ----

// One node's coefficients and their derivatives recomputed and contracted by AdjointSink.
// raw[g * 6L + ...] receives, per layer l (core first), t_x[l] = dH/dx_l and
// t_m[l] = sum conj(cot) da/dm_l (re, im) through the coefficients, and the absorption sum's
// derivatives [d/dx_l, d/dRe m_l, d/dIm m_l]: [t_x (L), t_m.re (L), t_m.im (L), jabs (3 L)].
// The solve holds most of the lane's registers, so the per-node gradient is assembled by
// ens_node_grads afterwards. Without table cotangents (abar null) the staged coefficients
// stand in for them with weight 0, so the order loop carries no branch on their presence.
template <int L>
__device__ __forceinline__ void vjp_node(const EnsArgs &a, int64_t e0, int64_t g,
                                         const int *node_n, const int64_t *node_off,
                                         const int64_t *ens_base, cx *slab, int64_t rows,
                                         cx *tile, const cx *staging, const cx *abar,
                                         const cx *backs, const double *sbar, double *raw) {
    const int64_t eb = g / a.n_radius, i = g % a.n_radius;
    const int64_t e = e0 + eb;
    const int64_t gi = e * a.n_radius + i;
    const NodeGeom geo = node_stored(a, e, i);
    const int n = node_n[gi];
    const int64_t off = ens_base[e] + node_off[gi];
    const double lam = a.wl[e], l2 = lam * lam;
    const double *sb = sbar + eb * 8;
    AdjointSink<L> sink;
    sink.init();
    sink.sa = staging + 2 * off;  // a_n, b_n interleaved
    sink.sb = staging + 2 * off + 1;
    const cx *tcot = abar ? abar : staging;
    sink.ta = tcot + 2 * off;
    sink.tb = tcot + 2 * off + n;
    sink.n = n;
    sink.tfac = abar ? l2 / kPi : 0.0;
    sink.alpha = (sb[1] + sb[2]) * l2 / (2.0 * kPi);
    sink.beta = sb[5] * l2 / (2.0 * kPi);
    sink.gamma = sb[3] * l2 / kPi;
    sink.back = backs[g];
    if constexpr (L == 1) {
        const cx m = mk(a.m[2 * e], a.m[2 * e + 1]);
        double *ux = real_column(slab, sphere_ckpt_rows(rows) * 64);
        const ZCheckpoint zs = ratio_chain_ckpt(m, geo.x, n, slab, 64, tile);
        x_chain(geo.x, n, upward_orders(geo.x, n), ux, 64);
        NoOuter none;
        mie_orders<true>(geo.x, m, n, zs, ux, 64, none, sink);
    } else {
        (void)tile;
        layered_solve<L, true>(geo.x, a.fr + e * (L - 1), a.m + 2 * e * L, n, slab, rows, sink);
    }
    double *o = raw + g * (6 * L);
#pragma unroll
    for (int l = 0; l < L; ++l) {
        o[l] = sink.tx[l];
        o[L + l] = sink.tm[l].re;
        o[2 * L + l] = sink.tm[l].im;
        o[3 * L + 3 * l] = sink.jabs[l][0];
        o[3 * L + 3 * l + 1] = sink.jabs[l][1];
        o[3 * L + 3 * l + 2] = sink.jabs[l][2];
    }
}

// Spheres: persistent over the ranked nodes, as ens_nodes_persist.
static __global__ __launch_bounds__(256) void ens_nodes_vjp_persist(
    EnsArgs a, int64_t e0, int64_t chunk_nodes, const int *order, const int *node_n,
    const int64_t *node_off, const int64_t *ens_base, unsigned long long *counter, cx *scratch,
    int64_t rows, const cx *staging, const cx *abar, const cx *backs, const double *sbar,
    double *raw) {
    __shared__ cx ztile[4][kZBlock * 64];
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    cx *tile = &ztile[threadIdx.x >> 6][lane];
    cx *slab = scratch + wave * slab_elems(1, rows) + lane;
    while (true) {
        unsigned long long chunk = 0;
        if (lane == 0) chunk = atomicAdd(counter, 1ull);
        chunk = __shfl(chunk, 0, 64);
        if ((int64_t)chunk * 64 >= chunk_nodes) break;
        const int64_t idx = (int64_t)chunk * 64 + lane;
        if (idx < chunk_nodes)
            vjp_node<1>(a, e0, order[idx], node_n, node_off, ens_base, slab, rows, tile, staging,
                        abar, backs, sbar, raw);
    }
}

// Layered spheres: one group of 64 ranked nodes per wavefront over [start, end), no work loop
// (as ens_nodes_ranked, for the same reason, with the same launches).
template <int L>
__global__ __launch_bounds__(256) void ens_nodes_vjp_ranked(
    EnsArgs a, int64_t e0, const int *order, int64_t start, int64_t end, const int *node_n,
    const int64_t *node_off, const int64_t *ens_base, cx *scratch, int64_t rows,
    const cx *staging, const cx *abar, const cx *backs, const double *sbar, double *raw) {
    __shared__ cx ztile[4][L == 1 ? kZBlock * 64 : 1];
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    cx *slab = scratch + wave * slab_elems(L, rows) + lane;
    cx *tile = L == 1 ? &ztile[threadIdx.x >> 6][lane] : nullptr;
    const int64_t idx = start + wave * 64 + lane;
    if (idx < end)
        vjp_node<L>(a, e0, order[idx], node_n, node_off, ens_base, slab, rows, tile, staging,
                    abar, backs, sbar, raw);
}

// One thread per node of the chunk: gnode[g * P + p], the node's share of the gradient,
// p = [p0, p1, p2, Re m_0, Im m_0, ..., Re m_{L-1}, Im m_{L-1}, f_0 .. f_{L-2}], from the
// solve's contractions and the node's weight and radius. dH/dx_l, dH/dRe m_l, dH/dIm m_l with
// H = phi / W; C_abs = (lambda^2/2pi) abs_sum enters C_ext and C_abs. With fixed_range the
// ln r range is held at its value (the distribution parameters act through the density alone,
// as the reference's derivatives do); otherwise the nodes and the range's half-width move with
// the parameters, the derivative of this discrete path. A monodisperse law's range is its
// radius, which moves its nodes either way.
static __global__ __launch_bounds__(256) void ens_node_grads(
    EnsArgs a, int64_t e0, int64_t chunk_nodes, const double *raw, const double *vals,
    const double *tabc, const double *ht, const double *sbar, int fixed_range, double *gnode) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= chunk_nodes) return;
    const int L = a.layers;
    const int P = 3 + 2 * L + (L - 1);
    const int64_t eb = g / a.n_radius, i = g % a.n_radius;
    const int64_t e = e0 + eb;
    const NodeGeom geo = node_stored(a, e, i);
    const double lam = a.wl[e], l2 = lam * lam;
    const double *sb = sbar + eb * 8;
    const double *rw = raw + g * (6 * L);
    const double *fr = a.fr + e * (L - 1);
    const double cabs = (sb[1] + sb[4]) * l2 / (2.0 * kPi);
    const double r = geo.r, W = geo.W;
    double *out = gnode + g * P;
    double dxt = 0.0;
    for (int l = 0; l < L; ++l) {
        const double dx = rw[l] + cabs * rw[3 * L + 3 * l];
        const double dre = rw[L + l] + cabs * rw[3 * L + 3 * l + 1];
        const double dim = -rw[2 * L + l] + cabs * rw[3 * L + 3 * l + 2];
        out[3 + 2 * l] = W * dre;
        out[4 + 2 * l] = W * dim;
        if (l < L - 1) {
            out[3 + 2 * L + l] = W * geo.x * dx;
            dxt += fr[l] * dx;
        } else {
            dxt += dx;
        }
    }
    const double dh_dlnr = geo.x * dxt + 2.0 * sb[6] * (r * r) + 3.0 * sb[7] * (r * r * r);
    const double *v = vals + g * kNodeVals;
    double phi = tabc[g] * ht[g];
    for (int k = 0; k < kNodeVals; ++k) phi += sb[k] * v[k];
    if (a.user_rw) {
        // The caller's nodes stay where they are: the parameters act through each node's
        // ln W alone, whose derivatives the caller gives (phi is the cotangent of ln W).
        const double *d = a.user_dlnw ? a.user_dlnw + (e * a.n_radius + i) * 3 : nullptr;
        for (int th = 0; th < 3; ++th) out[th] = d ? phi * d[th] : 0.0;
        return;
    }
    const int kind = a.kind[e];
    double p[3];
    ens_params(a, e, p);
    double dmid[3], dhalf[3], du[3];
    const double half = range_derivatives(kind, p, dmid, dhalf);
    log_density(kind, p, r, du);
    const double slope = log_density_slope(kind, p, r);
    const double xi = a.rnodes[i];
    const bool hold = fixed_range && kind != kMonodisperse;
    for (int th = 0; th < 3; ++th) {
        const double dlnr = hold ? 0.0 : dmid[th] + xi * dhalf[th];
        const double dlnw = (hold || half == 0.0 ? 0.0 : dhalf[th] / half) + du[th] +
                            (slope + 1.0) * dlnr;
        out[th] = phi * dlnw + W * dh_dlnr * dlnr;
    }
}

// grad[(e0 + eb) * P + p] = sum over the ensemble's nodes, in node order.
static __global__ void ens_grad_reduce(int64_t n_radius, int P, const double *gnode,
                                       double *grad) {
    const int64_t eb = blockIdx.x;
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        double s = 0.0;
        const double *gp = gnode + eb * n_radius * P + p;
        for (int64_t i = 0; i < n_radius; ++i) s += gp[i * P];
        grad[eb * P + p] = s;
    }
}

template <int L>
static void launch_nodes_vjp_ranked(hipStream_t s, const EnsArgs &a, int64_t e0,
                                    const EnsSetup &S, cx *scratch, const ChunkBufs &B,
                                    const cx *abar, const cx *backs, const double *sbar,
                                    double *raw) {
    for (const NodeLaunch &q : B.plan) {
        if constexpr (L == 1) {
            fill_u64<<<1, 1, 0, s>>>(S.counter, 0ull, 1);
            ens_nodes_vjp_persist<<<(unsigned)(q.waves / 4), 256, 0, s>>>(
                a, e0, q.end, B.order, S.node_n, S.node_off, S.ens_base, S.counter, scratch,
                q.rows, B.staging, abar, backs, sbar, raw);
        } else {
            ens_nodes_vjp_ranked<L><<<(unsigned)((q.end - q.start + 255) / 256), 256, 0, s>>>(
                a, e0, B.order, q.start, q.end, S.node_n, S.node_off, S.ens_base, scratch,
                q.rows, B.staging, abar, backs, sbar, raw);
        }
    }
}

}  // namespace scatrs

using namespace scatrs;

// The reverse pass of both entry points: nodes from the distributions, or with d_nodes the
// caller's (r, W) pairs and d_dlnw their d ln W / d p[3] (may be null: zero columns 0-2).
static int ensemble_reverse(scatrs_ctx *ctx, int64_t count, const int32_t *d_kind,
                            const double *d_p0, const double *d_p1, const double *d_p2,
                            const double *d_nodes, const double *d_dlnw,
                            const double *d_wavelength, int layers, const double *d_fractions,
                            const double *d_m, int64_t n_radius, int64_t n_angles, int64_t order,
                            int normalized, int fixed_range, int allow_unvalidated,
                            const double *d_bar_scalars, const double *d_bar_greek,
                            const double *d_bar_phase, double *d_grad, int64_t *total_terms) {
    if (!ctx) return SCATRS_INVALID_INPUT;
    ctx->error.clear();
    if (total_terms) *total_terms = 0;
    int rc = ens_check_args(ctx, count, layers, n_radius, n_angles, d_bar_greek != nullptr, order);
    if (rc) return rc;
    if (count == 0) return SCATRS_OK;
    if ((!d_nodes && (!d_kind || !d_p0 || !d_p1)) || !d_wavelength || !d_m || !d_grad ||
        (layers > 1 && !d_fractions))
        return fail(ctx, SCATRS_INVALID_INPUT, "null input or output pointer");
    EnsSetup S;
    if ((rc = ens_setup(ctx, count, d_kind, d_p0, d_p1, d_p2, d_wavelength, layers, d_fractions, d_m,
                        n_radius, n_angles, d_bar_greek != nullptr, order, allow_unvalidated,
                        ctx->angular_kernel, S, d_nodes)))
        return rc;
    S.a.user_dlnw = d_nodes ? d_dlnw : nullptr;
    // The reverse products read the FP64 coefficients: the node solves stage those, and with the
    // fp16-split kernel the forward products read an fp32 copy.
    S.stage32 = false;
    if (total_terms) *total_terms = S.total_terms;
    hipStream_t s = ctx->stream;
    const int P = 3 + 2 * layers + (layers - 1);
    const int h = S.h;
    const bool tables = d_bar_greek || d_bar_phase;
    // Phase cotangents at the caller's cosines (scatrs_set_phase_cosines): the forward tables
    // and the reverse products there too, the latter adding to the rule's (Greek cotangents).
    const bool user_mu = d_bar_phase && !ctx->phase_mu.empty();
    const bool rule_pass = d_bar_greek || (d_bar_phase && !user_mu);
    // Staging and cotangents take 32 B per term each (and the fp32 copy 16 B with kernel 2).
    const int64_t max_terms = chunk_terms_limit(6, 2 * sizeof(cx));
    const unsigned vjp_blocks = (unsigned)std::max(1, ctx->cus * 2);
    PhaseMu U;
    if (user_mu && (rc = phase_mu_setup(ctx, S, U))) return rc;
    const int n_out = user_mu ? 2 * U.h : 0;
    int64_t e0 = 0;
    while (e0 < count) {
        int64_t chunk_terms = 0;
        const int64_t e1 = next_chunk(S, e0, count, n_radius, max_terms, &chunk_terms);
        const int64_t ne = e1 - e0, nodes = ne * n_radius;
        if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kAdjNodes,
                                        (size_t)nodes * (3 + P + 6 * layers) * sizeof(double)),
                            "alloc")))
            return rc;
        double *ht = static_cast<double *>(ctx->buf[scatrs_ctx::kAdjNodes]);
        cx *backs = reinterpret_cast<cx *>(ht + nodes);
        double *gnode = ht + 3 * nodes;
        double *raw = gnode + P * nodes;
        ChunkBufs B;
        if ((rc = forward_nodes(ctx, S, e0, ne, chunk_terms, n_radius, n_angles, B, backs, s)))
            return rc;
        if (S.kernel == 2) {
            if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kStaging32,
                                            (size_t)std::max<int64_t>(chunk_terms, 1) * sizeof(float4)),
                                "staging")))
                return rc;
            float4 *st32 = static_cast<float4 *>(ctx->buf[scatrs_ctx::kStaging32]);
            stage_fp32<<<(unsigned)std::max(1, ctx->cus * 16), 256, 0, s>>>(B.staging, chunk_terms,
                                                                            st32);
            B.staging32 = st32;
        }
        if ((rc = forward_angular(ctx, S, e0, ne, n_radius, n_angles, B, s))) return rc;
        if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kAdjEns,
                                        (size_t)ne * (8 + 4 * n_angles + 8 * (size_t)n_out) *
                                            sizeof(double)),
                            "alloc")))
            return rc;
        double *sbar = static_cast<double *>(ctx->buf[scatrs_ctx::kAdjEns]);
        double *tbar = sbar + ne * 8;
        double *tab_u = tbar + ne * 4 * n_angles, *tbar_u = tab_u + ne * 4 * (size_t)n_out;
        if (user_mu &&
            (rc = angular_products(ctx, S, e0, ne, n_radius, U.h, n_out, U.table, U.tab16, B,
                                   tab_u, s)))
            return rc;
        const size_t shmem = (8 + 4 * (size_t)n_angles) * sizeof(double);
        if (d_bar_greek) {
            const dim3 grid((unsigned)((ne + 15) / 16), (unsigned)((n_angles + 63) / 64));
            ens_greek_vjp_mfma<<<grid, 256, 0, s>>>(e0, ne, (int)n_angles, S.d_atab, d_bar_greek,
                                                    S.d_dtab_t, order + 1, normalized, tbar);
        }
        ens_finish_vjp<<<(unsigned)ne, 256, shmem, s>>>(
            e0, n_radius, (int)n_angles, S.d_atab, B.vals, B.tab_out, S.d_dtab, order, normalized,
            d_bar_scalars, d_bar_greek, d_bar_phase, sbar, tbar, n_out, U.n_mu, tab_u, U.q_ptr,
            U.q_list, tbar_u);
        cx *abar = nullptr;
        if (tables) {
            if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kAdjStaging,
                                            (size_t)std::max<int64_t>(chunk_terms, 1) * 2 * sizeof(cx)),
                                "alloc")))
                return rc;
            const int hs = std::max(h, U.h);
            if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kAdjSlabs,
                                            (size_t)vjp_blocks * 2 * kVjpCols * hs * sizeof(double)),
                                "alloc")))
                return rc;
            abar = static_cast<cx *>(ctx->buf[scatrs_ctx::kAdjStaging]);
            double *slabs = static_cast<double *>(ctx->buf[scatrs_ctx::kAdjSlabs]);
            mark_begin(ctx);
            if (rule_pass) {
                fill_u64<<<1, 1, 0, s>>>(S.counter, 0ull, 1);
                ens_angular_vjp<<<vjp_blocks, 512, 0, s>>>(e0, ne, n_radius, (int)n_angles, h,
                                                           S.node_n, S.node_off, S.ens_base,
                                                           B.staging, S.table, tbar, S.counter,
                                                           slabs, abar, ht, 0);
            }
            if (user_mu) {
                fill_u64<<<1, 1, 0, s>>>(S.counter, 0ull, 1);
                ens_angular_vjp<<<vjp_blocks, 512, 0, s>>>(e0, ne, n_radius, n_out, U.h,
                                                           S.node_n, S.node_off, S.ens_base,
                                                           B.staging, U.table, tbar_u, S.counter,
                                                           slabs, abar, ht, rule_pass ? 1 : 0);
            }
            mark_end(ctx);
        } else {
            fill_f64<<<(unsigned)((nodes + 255) / 256), 256, 0, s>>>(ht, 0.0, nodes);
        }
        // The forward chunk's ranking and launch plan, and its scratch, serve the reverse solves:
        // the plan must be the slab layout's (a forward call's plan may instead solve small
        // launches one node per wavefront, which the reverse kernels cannot follow).
        for (const NodeLaunch &q : B.plan)
            if (q.waves < 0 || (layers == 1 && q.waves < 4))
                return fail(ctx, SCATRS_DEVICE_ERROR,
                            "internal: the reverse node solves met a launch plan without slabs");
        cx *scratch = static_cast<cx *>(ctx->buf[scatrs_ctx::kScratch]);
        const EnsArgs &a = S.a;
        mark_begin(ctx);
#define SCATRS_NODES_VJP(LL) \
    launch_nodes_vjp_ranked<LL>(s, a, e0, S, scratch, B, abar, backs, sbar, raw)
        switch (layers) {
            case 1: SCATRS_NODES_VJP(1); break;
            case 2: SCATRS_NODES_VJP(2); break;
            case 3: SCATRS_NODES_VJP(3); break;
            case 4: SCATRS_NODES_VJP(4); break;
            case 5: SCATRS_NODES_VJP(5); break;
            case 6: SCATRS_NODES_VJP(6); break;
            case 7: SCATRS_NODES_VJP(7); break;
            default: SCATRS_NODES_VJP(8); break;
        }
#undef SCATRS_NODES_VJP
        mark_end(ctx);
        ens_node_grads<<<(unsigned)((nodes + 255) / 256), 256, 0, s>>>(
            a, e0, nodes, raw, B.vals, B.tabc, ht, sbar, fixed_range, gnode);
        ens_grad_reduce<<<(unsigned)ne, 64, 0, s>>>(n_radius, P, gnode, d_grad + e0 * P);
        if ((rc = check_hip(ctx, hipGetLastError(), "ensemble vjp launch"))) return rc;
        e0 = e1;
    }
    return SCATRS_OK;
}

extern "C" int scatrs_ensemble_fixed_vjp(scatrs_ctx *ctx, int64_t count, const int32_t *d_kind,
                                         const double *d_p0, const double *d_p1,
                                         const double *d_p2, const double *d_wavelength,
                                         int layers,
                                         const double *d_fractions, const double *d_m,
                                         int64_t n_radius, int64_t n_angles, int64_t order,
                                         int normalized, int fixed_range,
                                         int allow_unvalidated, const double *d_bar_scalars,
                                         const double *d_bar_greek, const double *d_bar_phase,
                                         double *d_grad, int64_t *total_terms) {
    return ensemble_reverse(ctx, count, d_kind, d_p0, d_p1, d_p2, nullptr, nullptr, d_wavelength,
                            layers, d_fractions, d_m, n_radius, n_angles, order, normalized,
                            fixed_range, allow_unvalidated, d_bar_scalars, d_bar_greek,
                            d_bar_phase, d_grad, total_terms);
}

extern "C" int scatrs_ensemble_nodes_vjp(scatrs_ctx *ctx, int64_t count, int64_t n_nodes,
                                         const double *d_nodes, const double *d_dlnw,
                                         const double *d_wavelength, int layers,
                                         const double *d_fractions, const double *d_m,
                                         int64_t n_angles, int64_t order, int normalized,
                                         int allow_unvalidated, const double *d_bar_scalars,
                                         const double *d_bar_greek, const double *d_bar_phase,
                                         double *d_grad, int64_t *total_terms) {
    if (ctx && count > 0 && !d_nodes) {
        ctx->error.clear();
        return fail(ctx, SCATRS_INVALID_INPUT, "null input or output pointer");
    }
    return ensemble_reverse(ctx, count, nullptr, nullptr, nullptr, nullptr, d_nodes, d_dlnw,
                            d_wavelength, layers, d_fractions, d_m, n_nodes, n_angles, order,
                            normalized, 1, allow_unvalidated, d_bar_scalars, d_bar_greek,
                            d_bar_phase, d_grad, total_terms);
}
