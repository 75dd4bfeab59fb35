// icgn.hip -- per-POI inverse-compositional Gauss-Newton (IC-GN) for gfx950.
//
// Replaces ICGNSolver::solve_{first,second}_order_{znssd,ssd}[_masked]
// (src/subset/solver/icgn.cpp:446-1912).  Mapping: one 64-lane wavefront
// (one workgroup) per POI.  Lanes stride over the circular subset's samples
// (sample table in dy-outer/dx-inner order, icgn.cpp:485-489); every sum the
// reference forms serially is a per-lane partial followed by a butterfly
// (xor-shuffle) reduction, after which every lane holds the identical total.
//
// Once per POI the Hessian is accumulated (second order: from 3 x 15
// gradient-weighted monomial moments, since every entry is a coefficient
// times sum g_T dx^p dy^q with g_T in {gx^2, gx gy, gy^2}, p + q <= 4) and
// LDL^T-factored with Eigen 3.4 semantics (icgn.cpp:532-542; 12 x 12: one
// row per lane, 6 x 6: one lane), then turned into the solve operator H+:
// column j of H+ is the faithful LDLT solve of e_j (lane j computes it), so
// per iteration dp = -H+ G (icgn.cpp:594) is a lane-parallel mat-vec equal to
// the reference's solve up to rounding (the solve is linear in G, including
// its zeroing of vanishing pivots).  The inverse-compositional update runs on
// all lanes with uniform values: the affine one in closed form
// (icgn.cpp:83-120), the second-order one (icgn.cpp:289-357) as a
// wave-distributed full-pivot Gauss-Jordan elimination (one matrix entry per
// lane) with FullPivLU's pivot order and rank test.  Inside the iteration
// loop the reduced sums, dp and p pass through LDS; the workgroup is a
// single wave, so its barriers cost no waiting.
//
// Precision variants
//   R = double (plain drop-in): FP64 everywhere, the reference's two-pass
//     mean / norm / residual structure (icgn.cpp:577-592), deformed samples
//     kept in LDS between the passes.
//   R = float (mixed): FP32 coefficients, gradients, warp offsets (relative to
//     the integer part of u, v), sampling, products and per-lane sums; FP64
//     cross-lane (pairwise) reductions, Hessian factorisation, solve, warp
//     update and parameter accumulator.  ZNSSD sums are taken in one pass
//     about the shift K = mean(f):  sum(g-K), sum((g-K)^2), sum(fhat (g-K)),
//     sum((g-K) sd); mean, norm, G and C follow algebraically.
#include <cfloat>
#include <cmath>
#include <type_traits>

#include "icgn_common.hpp"

namespace dic {
namespace {

// One POI on the calling wave (64 lanes, the whole workgroup): sh is the
// workgroup's LDS record, gbuf (FP64 ZNSSD only) holds nfull doubles.
// Per-wave totals of a POI solved by NW > 1 waves (one workgroup), summed in
// wave order so that every wave holds the same totals.
template <int NW> struct CoopSlots {
    double sum[NW][16];
    double mom[NW][72]; // each wave's share of the reference-side moments
    int flag[NW];
    int bcast[2];
};

template <int P, int OBJ, int D, typename R, bool MASKED, int NW = 1>
__device__ __forceinline__ void solve_poi(const IcgnArgs &a, const int poi, IcgnShared &sh, double *gbuf,
                                          const Tile tile = Tile{nullptr, 0, 0, 0, 0},
                                          CoopSlots<NW> *cs = nullptr)
{
    constexpr bool ZNSSD = OBJ == DIC_OBJECTIVE_ZNSSD;
    constexpr bool FP64 = sizeof(R) == 8;
    constexpr bool TWOPASS = ZNSSD && FP64;
    // first-order FP32 kernels read a shared tile (they are bound by the load
    // data path; the second-order ones by the vector ALU)
    constexpr bool TILED = !FP64 && !MASKED && P == 6;

    const int lane = threadIdx.x & 63; // one POI per wave, or per workgroup of NW waves
    // NW > 1: wave 0 forms the reference-side sums, the factorisation, the
    // solve and the update; every wave takes a share of the samples in each
    // iteration, and the per-wave totals meet in cs
    const int wv = NW > 1 ? (int)(threadIdx.x >> 6) : 0;
    const bool leader = NW > 1 ? threadIdx.x == 0 : lane == 0;
    auto sync = [] {
        if constexpr (NW > 1) __syncthreads();
        else WAVE_SYNC();
    };
    // totals over the workgroup of n per-wave totals v[0..n) (uniform in each wave)
    auto coop_sum = [&](double *v, auto n) {
        if constexpr (NW > 1) {
            if (lane == 0) {
#pragma unroll
                for (int k = 0; k < decltype(n)::value; ++k) cs->sum[wv][k] = v[k];
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < decltype(n)::value; ++k) {
                double t = 0.0;
#pragma unroll
                for (int w = 0; w < NW; ++w) t += cs->sum[w][k];
                v[k] = t;
            }
            __syncthreads();
        }
    };
    const R *__restrict__ gxp = (const R *)a.gx;
    const R *__restrict__ gyp = (const R *)a.gy;
    const __amdgpu_buffer_rsrc_t crs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<void *>(a.coeffs), (short)0, a.coeff_bytes, 0x00020000);
    const int W = a.W, Hh = a.H;
    // FP32 kernels: reference, gradient planes and sample table through buffer
    // views (one 32-bit offset per sample for all three images; W * H * 4 is
    // below the coefficient image's size, itself < 2^31 bytes)
    const int pix_bytes = W * Hh * 4;
    const __amdgpu_buffer_rsrc_t grx = __builtin_amdgcn_make_buffer_rsrc(const_cast<R *>(gxp), (short)0, pix_bytes,
                                                                         0x00020000);
    const __amdgpu_buffer_rsrc_t gry = __builtin_amdgcn_make_buffer_rsrc(const_cast<R *>(gyp), (short)0, pix_bytes,
                                                                         0x00020000);
    const __amdgpu_buffer_rsrc_t frs = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(a.ref), (short)0,
                                                                         pix_bytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t trs = __builtin_amdgcn_make_buffer_rsrc(const_cast<int *>(a.table), (short)0,
                                                                         a.nfull * 4, 0x00020000);
    const double px = a.xy[2 * poi], py = a.xy[2 * poi + 1];
    const int cx = (int)round(px), cy = (int)round(py);

    // sample of the full circle used by this POI (masked: inside the image and the ROI)
    auto usable = [&](int x, int y) -> bool {
        if constexpr (MASKED) return x >= 0 && y >= 0 && x < W && y < Hh && a.roi[(size_t)y * W + x];
        return true;
    };

    int status = DIC_INVALID_INPUT;
    int iterations = 0;
    bool early = W <= 0 || Hh <= 0 || a.radius < 1 || a.max_it < 0;
    if (!early) {
        if constexpr (!MASKED) {
            early = cx - a.radius < 0 || cy - a.radius < 0 || cx + a.radius >= W || cy + a.radius >= Hh;
        } else {
            early = !usable(cx, cy);
        }
    }

    // ---- reference-side sums, once per POI (icgn.cpp:485-536) ----
    // Column walk of the circular subset: lane l takes subset column
    // dx = l - r + 64 g (column group g; one group for r <= 31) and runs down
    // it from the centre row, the samples at dy = +t and -t together (the disc
    // is symmetric about its centre row).  Every sum the solver needs is a
    // moment sum_{dx,dy} Q dx^p dy^q: the Hessian's with Q in {gx^2, gx gy, gy^2},
    // p + q <= 4 (P = 12) or 2, and the FP32 kernels' one-pass sums with Q in
    // {gx, gy, (f - mean) gx, (f - mean) gy}, p + q <= 2 or 1.  A lane forms the
    // column sums C_q = sum_dy Q dy^q (Q(+t) +- Q(-t) for even / odd q), the
    // moments are the lane sums of dx^p C_q, reduced a few dozen at a time by
    // wave_reduce_store (one walk for the Hessian moments and the norm, one for
    // the one-pass sums, so that few column sums are live at once).  FP64 in
    // both precisions: at small radii
    // these moments span many orders of magnitude (dx^4 against 1), and FP32
    // lane sums put errors into the ill-conditioned Hessian directions.
    constexpr int QH = (P == 12) ? 4 : 2, QG = (P == 12) ? 2 : 1; // largest p + q
    constexpr int NH = (QH + 1) * (QH + 2) / 2, NG = (QG + 1) * (QG + 2) / 2;
    constexpr bool REFSUMS = ZNSSD && !TWOPASS; // the FP32 kernels' one-pass ZNSSD sums
    // reduced sums in mred: [moments of gx^2 | norm | gx gy | gy^2 | sum gx, gy moments | sum (f - mean) |
    // sum (f - mean) gx, gy moments]
    constexpr int kNorm = NH, kRefG = 3 * NH + 1, kSumF = kRefG + 2 * NG, kRefF = kSumF + 1;
    auto hmom = [](int T, int k) { return T * NH + (T > 0 ? 1 : 0) + k; };
    const int ngroups = (2 * a.radius + 64) >> 6;
    // column of lane `lane` in group g and its half height h (-1: no column)
    auto column = [&](int g, int &dx, int &h) {
        dx = lane + 64 * g - a.radius;
        const int rem = a.radius * a.radius - dx * dx;
        int hh = rem < 0 ? -1 : (int)sqrtf((float)rem);
        if (rem >= 0 && (hh + 1) * (hh + 1) <= rem) ++hh;
        if (rem >= 0 && hh * hh > rem) --hh;
        h = hh;
    };
    // moment index k = d (d + 1) / 2 + q of the monomial dx^p dy^q, d = p + q
    auto mono_pq = [](int k, int &p, int &q) {
        int d = 0;
        while ((d + 1) * (d + 2) / 2 <= k) ++d;
        q = k - d * (d + 1) / 2;
        p = d - q;
    };
    double *mred = sh.Hinv; // reduced moments, until the solve operator is formed
    // NW > 1: wave w walks the rows t = w, w + NW, ... of every column and
    // reduces its share of the moments into cs->mom[w]
    double *wred = mred;
    if constexpr (NW > 1) wred = cs->mom[wv];
    if (!early) {
        double fs = 0.0;
        int cnt = 0;
        for (int g = 0; g < ngroups; ++g) {
            int dx, h;
            column(g, dx, h);
            const int x = cx + dx;
            for (int t = wv; t <= h; t += NW) {
                if (usable(x, cy + t)) {
                    fs += (double)a.ref[(size_t)(cy + t) * W + x];
                    ++cnt;
                }
                if (t > 0 && usable(x, cy - t)) {
                    fs += (double)a.ref[(size_t)(cy - t) * W + x];
                    ++cnt;
                }
            }
        }
        double fc[2] = {wave_sum(fs), (double)wave_sum_i(cnt)};
        coop_sum(fc, std::integral_constant<int, 2>{});
        const int ns = MASKED ? (int)fc[1] : a.nfull;
        if (ns < a.min_needed) {
            early = true;
        } else {
            const double fmean = ZNSSD ? fc[0] / (double)ns : 0.0;
            // column walk: calls visit(t, usable at +t, at -t, pixel offsets) per row pair
            auto walk = [&](int x, int h, auto &&visit) {
                for (int t = wv; t <= h; t += NW) {
                    const bool u1 = usable(x, cy + t), u2 = t > 0 && usable(x, cy - t);
                    visit((double)t, u1, u2, (size_t)(cy + t) * W + x, (size_t)(cy - t) * W + x);
                }
            };
            // c_q += (Q(+t) + Q(-t)) t^q for even q, (Q(+t) - Q(-t)) t^q for odd q
            auto acc = [](double *c, auto qmax, double t1, double q1, double q2) {
                const double sv = q1 + q2, dv = q1 - q2, t2 = t1 * t1;
                c[0] += sv;
                c[1] += dv * t1;
                if constexpr (decltype(qmax)::value >= 2) c[2] += sv * t2;
                if constexpr (decltype(qmax)::value >= 3) c[3] += dv * (t2 * t1);
                if constexpr (decltype(qmax)::value >= 4) c[4] += sv * (t2 * t2);
            };
            constexpr std::integral_constant<int, QH> qh{};
            constexpr std::integral_constant<int, QG> qg{};
            for (int g = 0; g < ngroups; ++g) {
                int dx, h;
                column(g, dx, h);
                const int x = cx + dx;
                const double d1 = (double)dx, d2 = d1 * d1;
                auto dxp = [&](int p) { return p == 0 ? 1.0 : (p == 1 ? d1 : (p == 2 ? d2 : (p == 3 ? d2 * d1 : d2 * d2))); };
                // moment k of column sums c: sum over lanes of dx^p c[q]
                auto moment = [&](const double *c, int k) {
                    int p, q;
                    mono_pq(k, p, q);
                    return dxp(p) * c[q];
                };
                {   // Hessian moments (and the norm), one reduction per gradient product
                    double cH[3][QH + 1], n2 = 0.0;
#pragma unroll
                    for (int q = 0; q <= QH; ++q) cH[0][q] = cH[1][q] = cH[2][q] = 0.0;
                    walk(x, h, [&](double t1, bool u1, bool u2, size_t o1, size_t o2) {
                        double gx1 = 0.0, gy1 = 0.0, gx2 = 0.0, gy2 = 0.0;
                        if (u1) {
                            gx1 = (double)gxp[o1];
                            gy1 = (double)gyp[o1];
                            if constexpr (ZNSSD) {
                                const double f = (double)a.ref[o1] - fmean;
                                n2 += f * f;
                            }
                        }
                        if (u2) {
                            gx2 = (double)gxp[o2];
                            gy2 = (double)gyp[o2];
                            if constexpr (ZNSSD) {
                                const double f = (double)a.ref[o2] - fmean;
                                n2 += f * f;
                            }
                        }
                        acc(cH[0], qh, t1, gx1 * gx1, gx2 * gx2);
                        acc(cH[1], qh, t1, gx1 * gy1, gx2 * gy2);
                        acc(cH[2], qh, t1, gy1 * gy1, gy2 * gy2);
                    });
                    wave_reduce_store<NH + 1>([&](int i) { return i < NH ? moment(cH[0], i) : n2; }, lane, wred,
                                              g > 0);
                    wave_reduce_store<NH>([&](int i) { return moment(cH[1], i); }, lane, wred + NH + 1, g > 0);
                    wave_reduce_store<NH>([&](int i) { return moment(cH[2], i); }, lane, wred + 2 * NH + 1, g > 0);
                }
                if constexpr (REFSUMS) { // sum sd, sum (f - mean) sd, sum (f - mean)
                    double cG[2][QG + 1], cF[2][QG + 1], sf = 0.0;
#pragma unroll
                    for (int q = 0; q <= QG; ++q) cG[0][q] = cG[1][q] = cF[0][q] = cF[1][q] = 0.0;
                    walk(x, h, [&](double t1, bool u1, bool u2, size_t o1, size_t o2) {
                        double gx1 = 0.0, gy1 = 0.0, f1 = 0.0, gx2 = 0.0, gy2 = 0.0, f2 = 0.0;
                        if (u1) {
                            gx1 = (double)gxp[o1];
                            gy1 = (double)gyp[o1];
                            f1 = (double)a.ref[o1] - fmean;
                        }
                        if (u2) {
                            gx2 = (double)gxp[o2];
                            gy2 = (double)gyp[o2];
                            f2 = (double)a.ref[o2] - fmean;
                        }
                        acc(cG[0], qg, t1, gx1, gx2);
                        acc(cG[1], qg, t1, gy1, gy2);
                        acc(cF[0], qg, t1, f1 * gx1, f2 * gx2);
                        acc(cF[1], qg, t1, f1 * gy1, f2 * gy2);
                        sf += f1 + f2;
                    });
                    wave_reduce_store<2 * NG + 1>(
                        [&](int i) { return i < NG ? moment(cG[0], i) : (i < 2 * NG ? moment(cG[1], i - NG) : sf); },
                        lane, wred + kRefG, g > 0);
                    wave_reduce_store<2 * NG>(
                        [&](int i) { return i < NG ? moment(cF[0], i) : moment(cF[1], i - NG); }, lane,
                        wred + kRefF, g > 0);
                }
            }
            if constexpr (NW > 1) { // the waves' shares, summed in wave order
                constexpr int kMoments = REFSUMS ? kRefF + 2 * NG : 3 * NH + 1;
                static_assert(kMoments <= 72, "moment slots");
                __syncthreads();
                if (wv == 0) {
                    for (int k = lane; k < kMoments; k += 64) {
                        double t = 0.0;
#pragma unroll
                        for (int w = 0; w < NW; ++w) t += cs->mom[w][k];
                        mred[k] = t;
                    }
                }
            }
            WAVE_SYNC();
            // the rest of the setup on wave 0 (its outcome reaches the other
            // waves after the factorisation)
            double fnorm = 1.0;
            if (ZNSSD && wv == 0) {
                fnorm = sqrt(mred[kNorm]);
                if (fnorm <= kEps) {
                    early = true;
                    status = DIC_NUMERICAL_FAILURE;
                }
            }
            if (leader) {
                sh.fmean = fmean;
                sh.fnorm = fnorm;
                sh.ns = ns;
            }
            if (!early && wv == 0) {
                // H = c * sum sd sd^T (icgn.cpp:532-536): entry (a, b) is c_a c_b times
                // the moment of g_t(a) g_t(b) dx^(p_a + p_b) dy^(q_a + q_b)
                const double hscale = ZNSSD ? 2.0 / (fnorm * fnorm) : 1.0;
                for (int e = lane; e < P * P; e += 64) {
                    const int ia = e / P, ib = e % P;
                    const ParamMonomial ma = param_monomial(ia), mb = param_monomial(ib);
                    const int qq = ma.q + mb.q, d = ma.p + mb.p + qq;
                    sh.H[e] = hscale * sd_half(ia) * sd_half(ib) * mred[hmom(ma.t + mb.t, d * (d + 1) / 2 + qq)];
                }
                if constexpr (REFSUMS) {
                    const double fin = 1.0 / fnorm;
                    if (lane < P) {
                        const ParamMonomial m = param_monomial(lane);
                        const int d = m.p + m.q, k = d * (d + 1) / 2 + m.q;
                        sh.ssd[lane] = sd_half(lane) * mred[kRefG + m.t * NG + k];
                        sh.sfsd[lane] = sd_half(lane) * fin * mred[kRefF + m.t * NG + k];
                    }
                    if (lane == 0) {
                        sh.sfh = mred[kSumF] * fin;
                        sh.sfh2 = mred[kNorm] * (fin * fin);
                    }
                }
            }
        }
    }

    if (!early && wv == 0) {
        WAVE_SYNC();
        int ok, pos;
        const bool fast = ldlt_operator_fast<P>(sh, lane, ok, pos);
        if (!fast) { // tied or NaN diagonal: Eigen's exchanges step by step
            if constexpr (P == 12) {
                ldlt_factor_wave<P>(sh, lane, ok, pos);
            } else { // 6 x 6: the serial form on one lane is as fast
                if (lane == 0) {
                    ldlt_factor<P>(sh.H, sh.tr, sh.x, &ok, &pos);
                    sh.state = ok && pos;
                }
                WAVE_SYNC();
                ok = pos = sh.state;
            }
        }
        if (!(ok && pos)) {
            early = true;
            status = DIC_NUMERICAL_FAILURE;
        } else if (!fast) {
            ldlt_operator_column<P>(sh, lane);
            WAVE_SYNC();
        }
    }
    if constexpr (NW > 1) { // wave 0's outcome to every wave
        if (leader) {
            cs->bcast[0] = early ? 1 : 0;
            cs->bcast[1] = status;
        }
        __syncthreads();
        early = cs->bcast[0] != 0;
        status = cs->bcast[1];
    }

    double corr = __builtin_inf();
    bool solved = false;
    if (!early) {
        if (leader) {
            const dic_initial ini = a.init[poi];
            sh.p[0] = ini.u;
            sh.p[1] = ini.v;
            sh.p[2] = ini.du_dx;
            sh.p[3] = ini.du_dy;
            sh.p[4] = ini.dv_dx;
            sh.p[5] = ini.dv_dy;
#pragma unroll
            for (int j = 6; j < 12; ++j) sh.p[j] = 0.0;
        }
        sync();
        const double Wm1 = (double)(W - 1), Hm1 = (double)(Hh - 1);
        const int ns = sh.ns;
        const double fmean = uniform_f64(sh.fmean), fnorm = uniform_f64(sh.fnorm);
        const R fm = (R)fmean, fin = (R)(1.0 / fnorm);
        bool converged = false, failed = false;
        for (int it = 0; it < a.max_it; ++it) {
            iterations = it + 1;
            double p[12];
#pragma unroll
            for (int j = 0; j < 12; ++j) p[j] = (j < P) ? uniform_f64(sh.p[j]) : 0.0;
            // a translation beyond any image puts every warped sample outside
            // (and would overflow the integer split below)
            const bool wild = !(fabs(p[0]) < 1e7 && fabs(p[1]) < 1e7);
            const double ui = wild ? 0.0 : floor(p[0]), vi = wild ? 0.0 : floor(p[1]);
            const R uf = (R)(p[0] - ui), vf = (R)(p[1] - vi);
            const int bx = cx + (int)ui, by = cy + (int)vi;
            R q[12];
#pragma unroll
            for (int j = 2; j < 12; ++j) q[j] = (R)(p[j] * sd_half(j)); // halves folded: q6, q8, q9, q11 meet x^2, y^2

            // whole warped subset provably inside [0, W-1] x [0, H-1]: the
            // per-sample bounds test can be skipped (the extent bounds the
            // linear part over the disc and the quadratic part by r^2 / 2)
            bool inside;
            bool staged = false;
            {
                const double rr = (double)a.radius;
                double ex = rr * sqrt((1.0 + p[2]) * (1.0 + p[2]) + p[3] * p[3]);
                double ey = rr * sqrt(p[4] * p[4] + (1.0 + p[5]) * (1.0 + p[5]));
                if constexpr (P == 12) {
                    ex += 0.5 * rr * rr * (fabs(p[6]) + fabs(p[7]) + fabs(p[8]));
                    ey += 0.5 * rr * rr * (fabs(p[9]) + fabs(p[10]) + fabs(p[11]));
                }
                // cx + p0 as bx + (p0 - ui): the same real number (p0 - ui is exact), rounded
                // once; built from this iteration's values, no double copy of the POI
                // centre is kept live across the loop (only read when !wild)
                const double x0 = (double)bx + (p[0] - ui), y0 = (double)by + (p[1] - vi), m = 0.01;
                inside = !wild && x0 - ex - m >= 0.0 && le_int(x0 + ex + m, W - 1) && y0 - ey - m >= 0.0 &&
                         le_int(y0 + ey + m, Hh - 1);
                if constexpr (TILED) {
                    // every sample's block inside the tile (the same extent bound)
                    if (inside && tile.t) {
                        const int xlo = (int)floor(x0 - ex - m), xhi = (int)floor(x0 + ex + m);
                        const int ylo = (int)floor(y0 - ey - m), yhi = (int)floor(y0 + ey + m);
                        staged = xlo >= tile.x0 && ylo >= tile.y0 && xhi + Spline<D>::n <= tile.x0 + tile.w &&
                                 yhi + Spline<D>::n <= tile.y0 + tile.h;
                    }
                }
            }
            bool oob = wild;
            Acc<R> s1, s2, sfg, sg[P];
            // the sample pass, instantiated once per source of the coefficients
            auto sample_pass = [&](auto from_tile) {
            // the next sample's table entry is loaded one sample ahead (reads
            // past the table return 0), so a sample waits for its gather only
            const int first = lane + 64 * wv;
            int tnext = (int)__builtin_amdgcn_raw_buffer_load_b32(trs, 4 * first, 0, 0);
            for (int s = wild ? a.nfull : first; s < a.nfull; s += 64 * NW) {
                int dx, dy;
                const int tcur = tnext;
                tnext = (int)__builtin_amdgcn_raw_buffer_load_b32(trs, 4 * (s + 64 * NW), 0, 0);
                decode(tcur, dx, dy);
                const int x = cx + dx, y = cy + dy;
                if (!usable(x, y)) continue;
                int ix, iy;
                R tx, ty;
                R xx = R(0), xy = R(0), yy = R(0); // FP32: monomials shared by the warp and the accumulation
                if constexpr (FP64) { // warp + bounds exactly as icgn.cpp:35-40, :559-575
                    const double lx = (double)dx, ly = (double)dy;
                    double wx = (double)cx + lx + p[0] + p[2] * lx + p[3] * ly;
                    double wy = (double)cy + ly + p[1] + p[4] * lx + p[5] * ly;
                    if constexpr (P == 12) {
                        const double dx22 = lx * lx * 0.5, dy22 = ly * ly * 0.5, dxy = lx * ly;
                        wx += p[6] * dx22 + p[7] * dxy + p[8] * dy22;
                        wy += p[9] * dx22 + p[10] * dxy + p[11] * dy22;
                    }
                    if (!inside && !(wx >= 0.0 && wy >= 0.0 && wx <= Wm1 && wy <= Hm1)) {
                        oob = true;
                        continue;
                    }
                    ix = min(max((int)floor(wx), 0), W - 1);
                    iy = min(max((int)floor(wy), 0), Hh - 1);
                    tx = (R)(wx - (double)ix);
                    ty = (R)(wy - (double)iy);
                } else { // FP32 offsets about the integer translation
                    const R lx = (R)dx, ly = (R)dy;
                    R xl = lx + uf + q[2] * lx + q[3] * ly;
                    R yl = ly + vf + q[4] * lx + q[5] * ly;
                    if constexpr (P == 12) {
                        xx = lx * lx;
                        xy = lx * ly;
                        yy = ly * ly;
                        xl += q[6] * xx + q[7] * xy + q[8] * yy;
                        yl += q[9] * xx + q[10] * xy + q[11] * yy;
                    }
                    // clamp before the int conversion; anything clamped is out of bounds
                    const R fx = fmin(fmax(floor(xl), R(-1e7)), R(1e7));
                    const R fy = fmin(fmax(floor(yl), R(-1e7)), R(1e7));
                    ix = bx + (int)fx;
                    iy = by + (int)fy;
                    tx = xl - fx;
                    ty = yl - fy;
                    // x' = ix + tx must lie in [0, W-1] (tx in [0, 1))
                    if (!inside && (ix < 0 || iy < 0 || ix > W - 1 || iy > Hh - 1 || (ix == W - 1 && tx > R(0)) ||
                                    (iy == Hh - 1 && ty > R(0)))) {
                        oob = true;
                        continue;
                    }
                }
                R gxv = R(0), gyv = R(0), fv = R(0);
                if constexpr (!TWOPASS) { // independent loads, issued before the gathers
                    if constexpr (FP64) {
                        gxv = gxp[(size_t)y * W + x];
                        gyv = gyp[(size_t)y * W + x];
                        fv = (R)a.ref[(size_t)y * W + x];
                    } else {
                        const int vo = (__mul24(y, W) + x) * 4;
                        gxv = buffer_f32(grx, vo);
                        gyv = buffer_f32(gry, vo);
                        fv = buffer_f32(frs, vo);
                    }
                }
                R g; // FP32: d!^2 times the B-spline value (unnormalised weights)
                if constexpr (FP64)
                    g = sample_at<D, R>(crs, a.pitch, a.b, ix, iy, tx, ty);
                else if constexpr (decltype(from_tile)::value)
                    g = sample_tile_u<D, kGroupPitch>(tile.t + (iy - tile.y0) * kGroupPitch + (ix - tile.x0), tx, ty);
                else
                    g = sample_at_u<D>(crs, a.pitch, a.b, ix, iy, tx, ty);
                if constexpr (TWOPASS) {
                    gbuf[s] = g;
                    s1.add(g);
                } else if constexpr (!FP64) {
                    const R lx = (R)dx, ly = (R)dy;
                    R w;
                    constexpr R us = SplineU<D>::inv_scale2;
                    if constexpr (ZNSSD) {
                        const R fh = (fv - fm) * fin;
                        w = g * us - fm;
                        s1.add(w);
                        s2.add(w * w);
                        sfg.add(fh * w);
                    } else { // SSD residual e = f - g, icgn.cpp:945-950
                        w = fv - g * us;
                        s2.add(w * w);
                    }
                    acc_sd<P, R>(sg, w * gxv, w * gyv, lx, ly, xx, xy, yy);
                } else {
                    R sd[P];
                    steepest_descent<P, R>(gxv, gyv, (R)dx, (R)dy, sd);
                    if constexpr (ZNSSD) {
                        const R fh = (fv - fm) * fin;
                        const R d = g - fm;
                        s1.add(d);
                        s2.add(d * d);
                        sfg.add(fh * d);
#pragma unroll
                        for (int j = 0; j < P; ++j) sg[j].add(d * sd[j]);
                    } else { // SSD residual e = f - g, icgn.cpp:945-950
                        const R e = fv - g;
                        s2.add(e * e);
#pragma unroll
                        for (int j = 0; j < P; ++j) sg[j].add(e * sd[j]);
                    }
                }
            }
            };
            if constexpr (TILED) {
                if (staged)
                    sample_pass(std::integral_constant<bool, true>{});
                else
                    sample_pass(std::integral_constant<bool, false>{});
            } else {
                sample_pass(std::integral_constant<bool, false>{});
            }
            bool any_oob = __any(oob);
            if constexpr (NW > 1) {
                if (lane == 0) cs->flag[wv] = any_oob ? 1 : 0;
                __syncthreads();
#pragma unroll
                for (int w = 0; w < NW; ++w) any_oob = any_oob || cs->flag[w] != 0;
                __syncthreads();
            }
            if (any_oob) {
                status = DIC_INVALID_INPUT;
                failed = true;
                break;
            }
            constexpr std::integral_constant<int, 1> one{};
            constexpr std::integral_constant<int, P + 1> g_and_c{};
            // ---- G and C into sh.red (lane 0 stores each wave total) ----
            if constexpr (TWOPASS) {
                // two-pass mean / norm / residual, icgn.cpp:577-592
                double t1[1] = {wave_sum(s1.value())};
                coop_sum(t1, one);
                const double gmean = t1[0] / (double)ns;
                double acc = 0.0;
                for (int s = lane + 64 * wv; s < a.nfull; s += 64 * NW) {
                    if constexpr (MASKED) {
                        int dx, dy;
                        decode(a.table[s], dx, dy);
                        if (!usable(cx + dx, cy + dy)) continue;
                    }
                    const double dg = gbuf[s] - gmean;
                    acc += dg * dg;
                }
                double t2[1] = {wave_sum(acc)};
                coop_sum(t2, one);
                const double gnorm = sqrt(t2[0]);
                if (gnorm <= kEps) {
                    status = DIC_NUMERICAL_FAILURE;
                    failed = true;
                    break;
                }
                double gl[P];
#pragma unroll
                for (int j = 0; j < P; ++j) gl[j] = 0.0;
                double cl = 0.0;
                for (int s = lane + 64 * wv; s < a.nfull; s += 64 * NW) {
                    int dx, dy;
                    decode(a.table[s], dx, dy);
                    const int x = cx + dx, y = cy + dy;
                    if (!usable(x, y)) continue;
                    double sd[P];
                    steepest_descent<P, double>((double)gxp[(size_t)y * W + x], (double)gyp[(size_t)y * W + x],
                                                (double)dx, (double)dy, sd);
                    const double fh = ((double)a.ref[(size_t)y * W + x] - fmean) / fnorm;
                    const double e = fh - (gbuf[s] - gmean) / gnorm;
#pragma unroll
                    for (int j = 0; j < P; ++j) gl[j] += e * sd[j];
                    cl += e * e;
                }
                const double sc = 2.0 / fnorm;
                double t[P + 1];
#pragma unroll
                for (int j = 0; j < P; ++j) t[j] = wave_sum(gl[j]);
                t[P] = wave_sum(cl);
                coop_sum(t, g_and_c);
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    const double v = t[j] * sc;
                    if (leader) sh.red[j] = v;
                }
                if (leader) sh.red[P] = t[P];
            } else if constexpr (ZNSSD) {
                // FP32 kernels: lane sums, then a pairwise FP32 tree over the wave (DPP)
                double t[P + 3];
#pragma unroll
                for (int j = 0; j < P; ++j) t[j] = (double)wave_total_f32(sg[j].s);
                t[P] = wave_total_f32(s1.s);
                t[P + 1] = wave_total_f32(s2.s);
                t[P + 2] = wave_total_f32(sfg.s);
                coop_sum(t, std::integral_constant<int, P + 3>{});
                const double S1 = t[P], S2 = t[P + 1], SFG = t[P + 2];
                const double md = S1 / (double)ns;
                const double var = S2 - S1 * md;
                if (!(var > 0.0) || sqrt(var) <= kEps) {
                    status = DIC_NUMERICAL_FAILURE;
                    failed = true;
                    break;
                }
                const double gnorm = sqrt(var), sc = 2.0 / fnorm, ig = 1.0 / gnorm;
#pragma unroll
                for (int j = 0; j < P; ++j) { // FP32 only: sg holds the monomial sums
                    const double v = t[j] * sd_half(j);
                    if (leader) sh.red[j] = sc * (sh.sfsd[j] - (v - md * sh.ssd[j]) * ig);
                }
                if (leader) sh.red[P] = sh.sfh2 - 2.0 * (SFG - md * sh.sfh) * ig + var * (ig * ig);
            } else {
                double t[P + 1];
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    if constexpr (FP64)
                        t[j] = wave_sum(sg[j].value());
                    else
                        t[j] = (double)wave_total_f32(sg[j].s) * sd_half(j);
                }
                if constexpr (FP64)
                    t[P] = wave_sum(s2.value());
                else
                    t[P] = (double)wave_total_f32(s2.s);
                coop_sum(t, g_and_c);
#pragma unroll
                for (int j = 0; j <= P; ++j)
                    if (leader) sh.red[j] = t[j];
            }
            sync();
            corr = sh.red[P];
            // ---- dp = -H+ G (icgn.cpp:594): lane i forms row i ----
            if (wv == 0 && lane < P) {
                double d = 0.0;
#pragma unroll
                for (int j = 0; j < P; ++j) d += sh.Hinv[lane * P + j] * sh.red[j];
                sh.dp[lane] = -d;
            }
            sync();
            bool fin = true;
            double nrm = 0.0;
#pragma unroll
            for (int j = 0; j < P; ++j) {
                const double d = sh.dp[j];
                fin = fin && isfinite(d);
                nrm += d * d;
            }
            if (!fin) {
                status = DIC_NUMERICAL_FAILURE;
                failed = true;
                break;
            }
            nrm = sqrt(nrm);
            bool invertible = true;
            if constexpr (P == 12) {
                if (wv == 0) invertible = second_order_solve_wave(sh, lane);
            }
            if (leader) {
                double pn[12];
                if constexpr (P == 6) {
                    ic_affine_update(sh.p, sh.dp, pn);
                } else if (invertible) {
                    // (X * P^-1): column 0 <- X(., 2), column 1 <- X(., 0), column 2 <- X(., 1)
                    pn[0] = sh.X[0][2];
                    pn[2] = sh.X[0][0] - 1.0;
                    pn[3] = sh.X[0][1];
                    pn[6] = 2.0 * sh.X[0][3];
                    pn[7] = sh.X[0][4];
                    pn[8] = 2.0 * sh.X[0][5];
                    pn[1] = sh.X[1][2];
                    pn[4] = sh.X[1][0];
                    pn[5] = sh.X[1][1] - 1.0;
                    pn[9] = 2.0 * sh.X[1][3];
                    pn[10] = sh.X[1][4];
                    pn[11] = 2.0 * sh.X[1][5];
                } else { // additive fallback, icgn.cpp:323-337
#pragma unroll
                    for (int j = 0; j < 12; ++j) pn[j] = sh.p[j] - sh.dp[j];
                }
                int ok = 1;
#pragma unroll
                for (int j = 0; j < P; ++j) {
                    sh.p[j] = pn[j];
                    ok = ok && isfinite(pn[j]);
                }
                sh.state = ok;
            }
            sync();
            if (!sh.state) {
                status = DIC_NUMERICAL_FAILURE;
                failed = true;
                break;
            }
            if (nrm < a.thr) { // icgn.cpp:607-610
                converged = true;
                break;
            }
        }
        if (!failed) {
            // running out of iterations still reports Success (icgn.cpp:620-622)
            if (!converged && isfinite(corr)) converged = true;
            status = converged ? DIC_SUCCESS : DIC_NOT_CONVERGED;
            solved = true;
        }
    }

    if (leader) {
        const dic_initial ini = a.init[poi];
        dic_result r;
        r.x = px;
        r.y = py;
#pragma unroll
        for (int j = 0; j < 12; ++j) r.p[j] = 0.0;
        if (solved) {
#pragma unroll
            for (int j = 0; j < P; ++j) r.p[j] = sh.p[j];
            r.correlation = corr;
        } else { // failure exits keep the initial u, v (icgn.cpp:455-462)
            r.p[0] = ini.u;
            r.p[1] = ini.v;
            r.correlation = ini.confidence;
        }
        r.status = status;
        r.valid = (status == DIC_SUCCESS) ? 1 : 0;
        r.iterations = iterations;
        r.reserved = 0;
        a.out[poi] = r;
    }
}

// Occupancy targets (waves per SIMD): mixed kernels 5 (at most 102 VGPRs;
// the second-order one spills 64 bytes per lane outside the sample loops and
// still runs ~4% faster than spill-free at 4 waves); FP64 second order 4
// (<= 128 VGPRs); FP64 first order takes what it needs.
template <int P, typename R> constexpr int kMinWaves = sizeof(R) == 4 ? 5 : (P == 12 ? 4 : 1);

template <int P, int OBJ, int D, typename R, bool MASKED>
__global__ __launch_bounds__(64, (kMinWaves<P, R>)) void icgn_kernel(IcgnArgs a)
{
    __shared__ IcgnShared sh;
    extern __shared__ double gbuf[];
    // workgroup k runs on XCD k % 8 (round-robin dispatch): each XCD takes one
    // contiguous eighth of the POIs, so neighbouring subsets share its L2
    const int per = (a.n + 7) >> 3;
    const int poi = (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
    if (poi >= a.n) return;
    solve_poi<P, OBJ, D, R, MASKED>(a, poi, sh, gbuf);
}

// A few POIs at a time (the heap-order propagation's launches): one POI per
// workgroup of kCoopWaves waves, which share its samples in every iteration,
// so that the launch takes about one POI's latency over kCoopWaves.
constexpr int kCoopWaves = 8;
template <int P, int OBJ, int D, typename R, bool MASKED>
__global__ __launch_bounds__(64 * kCoopWaves) void icgn_coop_kernel(IcgnArgs a)
{
    __shared__ IcgnShared sh;
    __shared__ CoopSlots<kCoopWaves> cs;
    extern __shared__ double gbuf[];
    const int poi = (int)blockIdx.x;
    if (poi >= a.n) return;
    solve_poi<P, OBJ, D, R, MASKED, kCoopWaves>(a, poi, sh, gbuf, Tile{nullptr, 0, 0, 0, 0}, &cs);
}

// ---- first-order FP32: kGroupPois neighbouring POIs per workgroup ----
// The workgroup stages the coefficient cells around the POIs' starting
// subsets (integer guess, strain up to r/10, 3 px margin) in LDS once; each
// wave then solves its POI on its own and reads the tile in every iteration
// whose warped subset lies inside it (the global gather otherwise), so the
// 36 reads of a sample come from LDS instead of the load data path.

// Grows the box [xlo, xhi] x [ylo, yhi] by POI poi's starting subset.
__device__ __forceinline__ void grow_box(const IcgnArgs &a, int poi, int &xlo, int &ylo, int &xhi, int &yhi)
{
    const dic_initial ini = a.init[poi];
    if (!(fabs(ini.u) < 1e6 && fabs(ini.v) < 1e6)) return;
    const int x = (int)round(a.xy[2 * poi]) + (int)floor(ini.u), y = (int)round(a.xy[2 * poi + 1]) + (int)floor(ini.v);
    const int e = a.radius + a.radius / 10 + 3;
    xlo = min(xlo, x - e);
    xhi = max(xhi, x + e);
    ylo = min(ylo, y - e);
    yhi = max(yhi, y + e);
}

// All threads of the workgroup: loads the box's cells into tile when it fits.
template <int D>
__device__ Tile stage_tile(const IcgnArgs &a, int xlo, int ylo, int xhi, int yhi, float *tile)
{
    using S = Spline<D>;
    xlo = max(xlo, 0);
    ylo = max(ylo, 0);
    xhi = min(xhi, a.W - 1);
    yhi = min(yhi, a.H - 1);
    const int tw = xhi - xlo + S::n, th = yhi - ylo + S::n;
    if (!(xhi >= xlo && yhi >= ylo && tw <= kGroupPitch && th <= kGroupRows)) return Tile{nullptr, 0, 0, 0, 0};
    const __amdgpu_buffer_rsrc_t crs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(a.coeffs), (short)0, a.coeff_bytes, 0x00020000);
    const int c0 = xlo + a.b - S::off, r0 = ylo + a.b - S::off;
    for (int e = threadIdx.x; e < kGroupPitch * th; e += blockDim.x) {
        const int row = e / kGroupPitch, col = e - row * kGroupPitch;
        if (col < tw) tile[e] = buffer_f32(crs, (__mul24(r0 + row, a.pitch) + c0 + col) * 4);
    }
    return Tile{tile, xlo, ylo, tw, th};
}

constexpr int kBigBox = 1 << 30;

template <int P, int OBJ, int D>
__global__ __launch_bounds__(64 * kGroupPois, (kMinWaves<P, float>)) void icgn_group_kernel(IcgnArgs a)
{
    __shared__ IcgnShared sh[kGroupPois];
    __shared__ float tile[kGroupPitch * kGroupRows];
    const int ngroups = (a.n + kGroupPois - 1) / kGroupPois;
    const int per = (ngroups + 7) >> 3; // XCD-blocked, as icgn_kernel
    const int g = (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
    if (g >= ngroups) return;
    int xlo = kBigBox, ylo = kBigBox, xhi = -kBigBox, yhi = -kBigBox;
    for (int q = 0; q < kGroupPois; ++q)
        if (g * kGroupPois + q < a.n) grow_box(a, g * kGroupPois + q, xlo, ylo, xhi, yhi);
    const Tile t = stage_tile<D>(a, xlo, ylo, xhi, yhi, tile);
    __syncthreads();
    const int poi = g * kGroupPois + (int)(threadIdx.x >> 6);
    if (poi >= a.n) return;
    solve_poi<P, OBJ, D, float, false>(a, poi, sh[threadIdx.x >> 6], nullptr, t);
}

// Serving ticks, first order FP32: POIs g .. g + 3 of the tick's contiguous
// ranges; the tile is staged when they all belong to one descriptor.
template <int OBJ, int D>
__global__ __launch_bounds__(64 * kGroupPois, (kMinWaves<6, float>)) void icgn_range_group_kernel(
    const BatchPair *__restrict__ pairs, const int *__restrict__ offsets, int npairs)
{
    __shared__ IcgnShared sh[kGroupPois];
    __shared__ float tile[kGroupPitch * kGroupRows];
    const int total = offsets[npairs];
    const int g0 = (int)blockIdx.x * kGroupPois;
    if (g0 >= total) return;
    auto pair_of = [&](int g) {
        int lo = 0, hi = npairs;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (offsets[mid] <= g) lo = mid;
            else hi = mid;
        }
        return lo;
    };
    const int gl = min(g0 + kGroupPois, total) - 1;
    const int p0 = pair_of(g0), pl = pair_of(gl);
    Tile t{nullptr, 0, 0, 0, 0};
    if (p0 == pl) {
        const IcgnArgs &a = pairs[p0].a;
        int xlo = kBigBox, ylo = kBigBox, xhi = -kBigBox, yhi = -kBigBox;
        for (int q = g0; q <= gl; ++q) grow_box(a, q - offsets[p0], xlo, ylo, xhi, yhi);
        t = stage_tile<D>(a, xlo, ylo, xhi, yhi, tile);
    }
    __syncthreads();
    const int g = g0 + (int)(threadIdx.x >> 6);
    if (g >= total) return;
    const int pid = __builtin_amdgcn_readfirstlane(pair_of(g));
    solve_poi<6, OBJ, D, float, false>(pairs[pid].a, g - offsets[pid], sh[threadIdx.x >> 6], nullptr,
                                       pid == p0 ? t : Tile{nullptr, 0, 0, 0, 0});
}

// Batched solving over many (pair, radius) descriptors: one launch per group
// of descriptors with the same shape order (each compiled like the per-pair
// kernel of that order; FP64 ZNSSD groups also share an LDS size class).
// items[] lists (descriptor, POI) in the order the host chose.
//   PERSISTENT = false: one workgroup (one wave) per item; the hardware
//     dispatcher hands out items in workgroup order as slots free up.
//   PERSISTENT = true: a fixed grid of workgroups takes `chunk` items at a
//     time from a queue counter (one atomic per chunk) until none are left.
// Each POI's result depends on nothing but its own inputs, so every mapping
// and order gives the results of the one-launch-per-pair kernels.
template <int P, int OBJ, int D, typename R>
__device__ __forceinline__ void run_item(const BatchPair *__restrict__ pairs, const int2 *__restrict__ items,
                                         unsigned u, IcgnShared &sh, double *gbuf)
{
    const int2 it = items[u];
    const int pid = __builtin_amdgcn_readfirstlane(it.x);
    const int poi = __builtin_amdgcn_readfirstlane(it.y);
    solve_poi<P, OBJ, D, R, false>(pairs[pid].a, poi, sh, gbuf);
}

template <int P, int OBJ, int D, typename R, bool PERSISTENT>
__global__ __launch_bounds__(64, (kMinWaves<P, R>)) void icgn_batch_kernel(
    const BatchPair *__restrict__ pairs, const int2 *__restrict__ items, int nitems, unsigned *counter, int chunk)
{
    __shared__ IcgnShared sh;
    extern __shared__ double gbuf[];
    if constexpr (!PERSISTENT) {
        if ((int)blockIdx.x < nitems) run_item<P, OBJ, D, R>(pairs, items, blockIdx.x, sh, gbuf);
    } else {
        for (;;) {
            unsigned t = 0;
            if (threadIdx.x == 0) t = atomicAdd(counter, (unsigned)chunk);
            t = (unsigned)__builtin_amdgcn_readfirstlane((int)__shfl(t, 0, 64));
            if (t >= (unsigned)nitems) break;
            const unsigned end = min(t + (unsigned)chunk, (unsigned)nitems);
            for (unsigned u = t; u < end; ++u) {
                __syncthreads();
                run_item<P, OBJ, D, R>(pairs, items, u, sh, gbuf);
            }
        }
    }
}

// Batch tile mapping, first order FP32: workgroup q takes items 4q .. 4q+3
// (quads dealt over the XCDs in contiguous runs, as icgn_group_kernel deals
// its POIs), one per wave; when the four belong to one descriptor the
// coefficient cells around their starting subsets are staged in LDS first.
// The item table keeps each descriptor's POIs in their natural order, so the
// four are neighbours.
template <int OBJ, int D>
__global__ __launch_bounds__(64 * kGroupPois, (kMinWaves<6, float>)) void icgn_batch_group_kernel(
    const BatchPair *__restrict__ pairs, const int2 *__restrict__ items, int nitems, unsigned *, int)
{
    __shared__ IcgnShared sh[kGroupPois];
    __shared__ float tile[kGroupPitch * kGroupRows];
    const int nq = (nitems + kGroupPois - 1) / kGroupPois;
    const int per = (nq + 7) >> 3;
    const int q = (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
    if (q >= nq) return;
    const int u0 = q * kGroupPois, ul = min(u0 + kGroupPois, nitems) - 1;
    const int p0 = items[u0].x;
    bool one = true;
    for (int u = u0 + 1; u <= ul; ++u) one = one && items[u].x == p0;
    Tile t{nullptr, 0, 0, 0, 0};
    if (one) {
        const IcgnArgs &a = pairs[p0].a;
        int xlo = kBigBox, ylo = kBigBox, xhi = -kBigBox, yhi = -kBigBox;
        for (int u = u0; u <= ul; ++u) grow_box(a, items[u].y, xlo, ylo, xhi, yhi);
        t = stage_tile<D>(a, xlo, ylo, xhi, yhi, tile);
    }
    __syncthreads();
    const int u = u0 + (int)(threadIdx.x >> 6);
    if (u > ul) return;
    const int2 it = items[u];
    const int pid = __builtin_amdgcn_readfirstlane(it.x);
    const int poi = __builtin_amdgcn_readfirstlane(it.y);
    solve_poi<6, OBJ, D, float, false>(pairs[pid].a, poi, sh[threadIdx.x >> 6], nullptr, t);
}

// Serving ticks: descriptors own contiguous POI ranges [offsets[k],
// offsets[k+1]); workgroup g finds its descriptor by a binary search over
// the (npairs + 1) prefix offsets, so a tick needs no per-POI item table.
template <int P, int OBJ, int D, typename R>
__global__ __launch_bounds__(64, (kMinWaves<P, R>)) void icgn_range_kernel(
    const BatchPair *__restrict__ pairs, const int *__restrict__ offsets, int npairs)
{
    __shared__ IcgnShared sh;
    extern __shared__ double gbuf[];
    const int g = (int)blockIdx.x;
    if (g >= offsets[npairs]) return;
    int lo = 0, hi = npairs;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (offsets[mid] <= g) lo = mid;
        else hi = mid;
    }
    const int pid = __builtin_amdgcn_readfirstlane(lo);
    solve_poi<P, OBJ, D, R, false>(pairs[pid].a, g - offsets[pid], sh, gbuf);
}

// Next frame's initial guesses from this frame's results (u, v, the four
// displacement gradients and the correlation as confidence).
__global__ void init_from_results_kernel(const dic_result *__restrict__ prev, dic_initial *__restrict__ init, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const dic_result r = prev[i];
    dic_initial o;
    o.u = r.p[0];
    o.v = r.p[1];
    o.du_dx = r.p[2];
    o.du_dy = r.p[3];
    o.dv_dx = r.p[4];
    o.dv_dy = r.p[5];
    o.confidence = r.correlation;
    init[i] = o;
}

// Next frame's guesses extrapolated along the sequence: a = results K frames
// back, b = 2K frames back; u, v and the four gradients of 2 a - b where both
// records are Success, a's values otherwise; confidence = a's correlation.
__global__ void init_extrapolated_kernel(const dic_result *__restrict__ a, const dic_result *__restrict__ b,
                                         dic_initial *__restrict__ init, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const dic_result ra = a[i], rb = b[i];
    const bool lin = ra.status == DIC_SUCCESS && rb.status == DIC_SUCCESS;
    double q[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) q[j] = lin ? 2.0 * ra.p[j] - rb.p[j] : ra.p[j];
    init[i] = dic_initial{q[0], q[1], q[2], q[3], q[4], q[5], ra.correlation};
}

__global__ void fill_u32_kernel(unsigned *p, int n, unsigned v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Forward-compositional Gauss-Newton (ForwardGaussNewtonSolver,
// src/subset/solver/forward_gauss_newton.cpp).  Per iteration the deformed
// image is sampled with its B-spline value and gradient at every warped
// point; the steepest-descent vector is that gradient through dW/dz of the
// current warp (:44-60, :986-1019); H, G are rebuilt, LDLT-factored, and
// dp = +H^-1 G is capped and composed as W(p) o W(dp) (:64-76, :1025-1053).
// One pass over the samples per iteration, nothing stored per sample: every
// quantity the iteration needs is a sum over the subset, accumulated per lane
// and reduced once.  The Hessian is 3 x NM moments of {ge^2, ge he, he^2}
// times the monomials dx^p dy^q (ge, he: the projected gradient); G and C
// of ZNSSD come from the sums of d = g - mean_f, d^2, f^ d and of f^ sd,
// d sd, sd (f^ the normalised reference):
//   e = f^ - (d - mean_d) / |d - mean_d|,  G = 2 / |f - mean_f| sum e sd
//     = 2 / |f - mean_f| (sum f^ sd - (sum d sd - mean_d sum sd) / |d - mean_d|),
//   C = sum f^2 - 2 (sum f^ d - mean_d sum f^) / |d - mean_d| + 1,
// the same quantities as the reference's two-pass forms (:31-39, :335-350),
// rounded differently.  The solve is the LDLT operator of the IC-GN setup,
// formed per iteration.  Variant constants as the reference has them: step
// cap 0.5 / 0.35 / 0.3, which norm the convergence test reads,
// Levenberg-Marquardt on three of the four second-order variants.
// ---------------------------------------------------------------------------
template <int P, int OBJ, bool MASKED> struct FgnRules {
    static constexpr bool znssd = OBJ == DIC_OBJECTIVE_ZNSSD;
    // :367, :573, :763, :936, :1260, :1485, :1700, :1902
    static constexpr double cap = P == 6 ? (znssd ? 0.5 : 0.35) : (znssd ? (MASKED ? 0.3 : 0.5) : 0.35);
    // the convergence test reads the norm before the cap (:763-767, :1485-1489)
    static constexpr bool precap = (P == 6 && !znssd && !MASKED) || (P == 12 && znssd && MASKED);
    // H_ii += 1e-4 max_i H_ii (:1242-1245, :1682-1685, :1886-1889)
    static constexpr bool lm = P == 12 && !(znssd && MASKED);
};

// compose_warp: W(x; pn) = W(x; p) o W(x; dp) (:64-76 first order,
// :1025-1053 second order, first order in dp)
template <int P> __device__ void fgn_compose(const double *p, const double *dp, double *pn)
{
    pn[0] = dp[0] + p[0] + p[2] * dp[0] + p[3] * dp[1];
    pn[1] = dp[1] + p[1] + p[4] * dp[0] + p[5] * dp[1];
    if constexpr (P == 6) {
        pn[2] = dp[2] + p[2] * (1.0 + dp[2]) + p[3] * dp[4];
        pn[3] = dp[3] + p[2] * dp[3] + p[3] * (1.0 + dp[5]);
        pn[4] = dp[4] + p[4] * (1.0 + dp[2]) + p[5] * dp[4];
        pn[5] = dp[5] + p[4] * dp[3] + p[5] * (1.0 + dp[5]);
    } else {
        pn[2] = dp[2] + p[2] * (1.0 + dp[2]) + p[3] * dp[4] + p[6] * dp[0] + p[7] * dp[1];
        pn[3] = dp[3] + p[2] * dp[3] + p[3] * (1.0 + dp[5]) + p[7] * dp[0] + p[8] * dp[1];
        pn[4] = dp[4] + p[4] * (1.0 + dp[2]) + p[5] * dp[4] + p[9] * dp[0] + p[10] * dp[1];
        pn[5] = dp[5] + p[4] * dp[3] + p[5] * (1.0 + dp[5]) + p[10] * dp[0] + p[11] * dp[1];
        pn[6] = dp[6] + p[2] * dp[6] + p[3] * dp[9] + p[6] * (1.0 + 2.0 * dp[2]) + 2.0 * p[7] * dp[4];
        pn[7] = dp[7] + p[2] * dp[7] + p[3] * dp[10] + p[6] * dp[3] + p[7] * (1.0 + dp[2] + dp[5]) + p[8] * dp[4];
        pn[8] = dp[8] + p[2] * dp[8] + p[3] * dp[11] + p[7] * dp[3] + p[8] * (1.0 + 2.0 * dp[5]);
        pn[9] = dp[9] + p[4] * dp[6] + p[5] * dp[9] + p[9] * (1.0 + 2.0 * dp[2]) + 2.0 * p[10] * dp[4];
        pn[10] = dp[10] + p[4] * dp[7] + p[5] * dp[10] + p[9] * dp[3] + p[10] * (1.0 + dp[2] + dp[5]) +
                 p[11] * dp[4];
        pn[11] = dp[11] + p[4] * dp[8] + p[5] * dp[11] + p[10] * dp[3] + p[11] * (1.0 + 2.0 * dp[5]);
    }
}

// B-spline value and gradient at (ix + tx, iy + ty) (BSplineInterpolator::value
// and ::gradient, bspline.cpp:458-500, as a tensor product)
template <int D, typename R>
__device__ __forceinline__ void sample_vg(__amdgpu_buffer_rsrc_t rs, int pitch, int b, int ix, int iy, R tx, R ty,
                                          R &v, R &gx, R &gy)
{
    using S = Spline<D>;
    const int voff = (__mul24(iy + b - S::off, pitch) + (ix + b - S::off)) * (int)sizeof(R);
    const int rowb = pitch * (int)sizeof(R);
    R wx[S::n], wy[S::n], dwx[S::n], dwy[S::n];
    S::weights(tx, wx);
    S::weights(ty, wy);
    S::dweights(tx, dwx);
    S::dweights(ty, dwy);
    v = gx = gy = R(0);
#pragma unroll
    for (int j = 0; j < S::n; ++j) {
        R c[S::n];
        RowLoad<R>::template row<S::n>(rs, voff, j * rowb, c);
        R rw = R(0), rd = R(0);
#pragma unroll
        for (int k = 0; k < S::n; ++k) {
            rw += wx[k] * c[k];
            rd += dwx[k] * c[k];
        }
        v += wy[j] * rw;
        gx += wy[j] * rd;
        gy += dwy[j] * rw;
    }
}

// Occupancy targets of the FGN kernels (waves per SIMD), set by the per-lane
// sums they keep: 3 x NM Hessian moments plus 3P + 3 (ZNSSD) or P + 1 (SSD).
template <int P, typename R> constexpr int kFgnMinWaves = sizeof(R) == 4 ? (P == 12 ? 3 : 5) : (P == 12 ? 2 : 3);

template <int P, int OBJ, int D, typename R, bool MASKED>
__device__ __forceinline__ void fgn_solve_poi(const IcgnArgs &a, const int poi, IcgnShared &sh)
{
    constexpr bool ZNSSD = OBJ == DIC_OBJECTIVE_ZNSSD;
    constexpr bool FP64 = sizeof(R) == 8;
    using Rules = FgnRules<P, OBJ, MASKED>;
    constexpr int NM = P == 12 ? 15 : 6; // monomials dx^p dy^q with p + q <= 2 (P = 6) or 4 (P = 12)
    const int lane = threadIdx.x;
    const __amdgpu_buffer_rsrc_t crs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<void *>(a.coeffs), (short)0, a.coeff_bytes, 0x00020000);
    const int W = a.W, Hh = a.H;
    const double px = a.xy[2 * poi], py = a.xy[2 * poi + 1];
    const int cx = (int)round(px), cy = (int)round(py);
    auto usable = [&](int x, int y) -> bool {
        if constexpr (MASKED) return x >= 0 && y >= 0 && x < W && y < Hh && a.roi[(size_t)y * W + x];
        return true;
    };

    int status = DIC_INVALID_INPUT;
    int iterations = 0;
    bool early = W <= 0 || Hh <= 0 || a.radius < 1 || a.max_it < 0;
    if (!early) {
        if constexpr (!MASKED) early = cx - a.radius < 0 || cy - a.radius < 0 || cx + a.radius >= W ||
                                       cy + a.radius >= Hh;
        else early = !usable(cx, cy);
    }
    // reference statistics (forward_gauss_newton.cpp:233-274), and for ZNSSD
    // the sums of f^ and f^2 in the precision the iterations form f^ in
    if (!early) {
        double fsum_l = 0.0;
        int cnt = 0;
        for (int s = lane; s < a.nfull; s += 64) {
            int dx, dy;
            decode(a.table[s], dx, dy);
            if (!usable(cx + dx, cy + dy)) continue;
            fsum_l += (double)a.ref[(size_t)(cy + dy) * W + cx + dx];
            ++cnt;
        }
        const int ns = wave_sum_i(cnt);
        double fmean = 0.0, fnorm = 1.0, sfh = 0.0, sfh2 = 0.0;
        if (ns < a.min_needed) {
            early = true;
        } else if constexpr (ZNSSD) {
            fmean = wave_sum(fsum_l) / (double)ns;
            double acc = 0.0;
            for (int s = lane; s < a.nfull; s += 64) {
                int dx, dy;
                decode(a.table[s], dx, dy);
                if (!usable(cx + dx, cy + dy)) continue;
                const double df = (double)a.ref[(size_t)(cy + dy) * W + cx + dx] - fmean;
                acc += df * df;
            }
            fnorm = sqrt(wave_sum(acc));
            if (fnorm <= kEps) {
                early = true;
                status = DIC_NUMERICAL_FAILURE;
            } else {
                const R fm = (R)fmean, fin = (R)(1.0 / fnorm);
                double h1 = 0.0, h2 = 0.0;
                for (int s = lane; s < a.nfull; s += 64) {
                    int dx, dy;
                    decode(a.table[s], dx, dy);
                    if (!usable(cx + dx, cy + dy)) continue;
                    const double fh = (double)(((R)a.ref[(size_t)(cy + dy) * W + cx + dx] - fm) * fin);
                    h1 += fh;
                    h2 += fh * fh;
                }
                sfh = wave_sum(h1);
                sfh2 = wave_sum(h2);
            }
        }
        if (lane == 0) {
            sh.fmean = fmean;
            sh.fnorm = fnorm;
            sh.sfh = sfh;
            sh.sfh2 = sfh2;
            sh.ns = ns;
        }
    }

    double corr = __builtin_inf();
    bool solved = false;
    if (!early) {
        if (lane == 0) {
            const dic_initial ini = a.init[poi];
            sh.p[0] = ini.u;
            sh.p[1] = ini.v;
            sh.p[2] = ini.du_dx;
            sh.p[3] = ini.du_dy;
            sh.p[4] = ini.dv_dx;
            sh.p[5] = ini.dv_dy;
#pragma unroll
            for (int j = 6; j < 12; ++j) sh.p[j] = 0.0;
        }
        WAVE_SYNC();
        const double Wm1 = (double)(W - 1), Hm1 = (double)(Hh - 1);
        const int ns = sh.ns;
        const double fmean = uniform_f64(sh.fmean), fnorm = uniform_f64(sh.fnorm);
        const R fm = (R)fmean, fin = (R)(1.0 / fnorm);
        // raw sums of the iteration, reduced into the operator's space (the
        // operator is formed only after they are read)
        double *const raw = sh.Hinv;
        bool converged = false, failed = false;
        for (int it = 0; it < a.max_it; ++it) {
            iterations = it + 1;
            double p[12];
#pragma unroll
            for (int j = 0; j < 12; ++j) p[j] = (j < P) ? uniform_f64(sh.p[j]) : 0.0;
            const bool wild = !(fabs(p[0]) < 1e7 && fabs(p[1]) < 1e7);
            const double ui = wild ? 0.0 : floor(p[0]), vi = wild ? 0.0 : floor(p[1]);
            const R uf = (R)(p[0] - ui), vf = (R)(p[1] - vi);
            const int bx = cx + (int)ui, by = cy + (int)vi;
            R q[12];
#pragma unroll
            for (int j = 2; j < 12; ++j) q[j] = (R)p[j];
            bool oob = wild;
            // ZNSSD: s1 = sum d, s2 = sum d^2, s3 = sum f^ d; ga = sum f^ sd,
            // gb = sum d sd, gc = sum sd.  SSD: s2 = sum e^2, ga = sum e sd.
            Acc<R> s1, s2, s3, ga[P], gb[P], gc[P], hm[3 * NM];
            for (int s = wild ? a.nfull : lane; s < a.nfull; s += 64) {
                int dx, dy;
                decode(a.table[s], dx, dy);
                const int x = cx + dx, y = cy + dy;
                if (!usable(x, y)) continue;
                int ix, iy;
                R tx, ty;
                const R lx = (R)dx, ly = (R)dy;
                // warp and bounds (:283-300)
                if constexpr (FP64) {
                    double wx = (double)cx + lx + p[0] + p[2] * lx + p[3] * ly;
                    double wy = (double)cy + ly + p[1] + p[4] * lx + p[5] * ly;
                    if constexpr (P == 12) {
                        const double dx22 = lx * lx * 0.5, dy22 = ly * ly * 0.5, dxy = lx * ly;
                        wx += p[6] * dx22 + p[7] * dxy + p[8] * dy22;
                        wy += p[9] * dx22 + p[10] * dxy + p[11] * dy22;
                    }
                    if (!(wx >= 0.0 && wy >= 0.0 && wx <= Wm1 && wy <= Hm1)) {
                        oob = true;
                        continue;
                    }
                    ix = min(max((int)floor(wx), 0), W - 1);
                    iy = min(max((int)floor(wy), 0), Hh - 1);
                    tx = (R)(wx - (double)ix);
                    ty = (R)(wy - (double)iy);
                } else {
                    R xl = lx + uf + q[2] * lx + q[3] * ly;
                    R yl = ly + vf + q[4] * lx + q[5] * ly;
                    if constexpr (P == 12) {
                        const R dx22 = lx * lx * R(0.5), dy22 = ly * ly * R(0.5), dxy = lx * ly;
                        xl += q[6] * dx22 + q[7] * dxy + q[8] * dy22;
                        yl += q[9] * dx22 + q[10] * dxy + q[11] * dy22;
                    }
                    const R fx = fmin(fmax(floor(xl), R(-1e7)), R(1e7));
                    const R fy = fmin(fmax(floor(yl), R(-1e7)), R(1e7));
                    ix = bx + (int)fx;
                    iy = by + (int)fy;
                    tx = xl - fx;
                    ty = yl - fy;
                    if (ix < 0 || iy < 0 || ix > W - 1 || iy > Hh - 1 || (ix == W - 1 && tx > R(0)) ||
                        (iy == Hh - 1 && ty > R(0))) {
                        oob = true;
                        continue;
                    }
                }
                const R f = (R)a.ref[(size_t)y * W + x];
                R g, gx, gy;
                sample_vg<D, R>(crs, a.pitch, a.b, ix, iy, tx, ty, g, gx, gy);
                // dW/dz of the current warp at (lx, ly) (:994-998)
                R a00 = R(1) + q[2], a01 = q[3], a10 = q[4], a11 = R(1) + q[5];
                if constexpr (P == 12) {
                    a00 += q[6] * lx + q[7] * ly;
                    a01 += q[7] * lx + q[8] * ly;
                    a10 += q[9] * lx + q[10] * ly;
                    a11 += q[10] * lx + q[11] * ly;
                }
                const R ge = gx * a00 + gy * a10, he = gx * a01 + gy * a11;
                const R xx = lx * lx, xy = lx * ly, yy = ly * ly;
                if constexpr (ZNSSD) {
                    const R fh = (f - fm) * fin, d = g - fm;
                    s1.add(d);
                    s2.add(d * d);
                    s3.add(fh * d);
                    acc_sd<P, R>(ga, fh * ge, fh * he, lx, ly, xx, xy, yy);
                    acc_sd<P, R>(gb, d * ge, d * he, lx, ly, xx, xy, yy);
                    acc_sd<P, R>(gc, ge, he, lx, ly, xx, xy, yy);
                } else { // e = f - g
                    const R e = f - g;
                    s2.add(e * e);
                    acc_sd<P, R>(ga, e * ge, e * he, lx, ly, xx, xy, yy);
                }
                const R Q[3] = {ge * ge, ge * he, he * he};
                R mono[NM];
                mono[0] = R(1);
                mono[1] = lx;
                mono[2] = ly;
                mono[3] = xx;
                mono[4] = xy;
                mono[5] = yy;
                if constexpr (P == 12) {
                    mono[6] = xx * lx;
                    mono[7] = xx * ly;
                    mono[8] = lx * yy;
                    mono[9] = yy * ly;
                    mono[10] = xx * xx;
                    mono[11] = xx * xy;
                    mono[12] = xx * yy;
                    mono[13] = xy * yy;
                    mono[14] = yy * yy;
                }
#pragma unroll
                for (int T = 0; T < 3; ++T)
#pragma unroll
                    for (int k = 0; k < NM; ++k) hm[T * NM + k].add(Q[T] * mono[k]);
            }
            if (__any(oob)) {
                status = DIC_INVALID_INPUT;
                failed = true;
                break;
            }
            // ---- the sums over the wave: Hessian moments, then G and C ----
#pragma unroll
            for (int T = 0; T < 3; ++T)
                wave_reduce_store<NM>([&](int i) { return hm[T * NM + i].value(); }, lane, sh.mom[T], false);
            if constexpr (ZNSSD) {
                wave_reduce_store<3 + P>(
                    [&](int i) { return i == 0 ? s1.value() : (i == 1 ? s2.value() : (i == 2 ? s3.value() : ga[i - 3].value())); },
                    lane, raw, false);
                wave_reduce_store<P>([&](int i) { return gb[i].value(); }, lane, raw + 3 + P, false);
                wave_reduce_store<P>([&](int i) { return gc[i].value(); }, lane, raw + 3 + 2 * P, false);
            } else {
                wave_reduce_store<1 + P>([&](int i) { return i == 0 ? s2.value() : ga[i - 1].value(); }, lane, raw,
                                         false);
            }
            WAVE_SYNC();
            if constexpr (ZNSSD) {
                // deformed mean and norm (vector_norm, :31-39) from the shifted sums
                const double S1 = raw[0], md = S1 / (double)ns, var = raw[1] - S1 * md;
                const double gnorm = sqrt(var);
                if (!(var > 0.0) || gnorm <= kEps) {
                    status = DIC_NUMERICAL_FAILURE;
                    failed = true;
                    break;
                }
                const double ig = 1.0 / gnorm, gsc = 2.0 / fnorm;
                if (lane < P)
                    sh.red[lane] = gsc * sd_half(lane) *
                                   (raw[3 + lane] - (raw[3 + P + lane] - md * raw[3 + 2 * P + lane]) * ig);
                if (lane == 0) sh.red[P] = sh.sfh2 - 2.0 * (raw[2] - md * sh.sfh) * ig + var * (ig * ig);
            } else {
                if (lane < P) sh.red[lane] = sd_half(lane) * raw[1 + lane];
                if (lane == 0) sh.red[P] = raw[0];
            }
            {
                const double hsc = ZNSSD ? 2.0 / (fnorm * fnorm) : 1.0;
                for (int e = lane; e < P * P; e += 64) {
                    const int ia = e / P, ib = e % P;
                    const ParamMonomial ma = param_monomial(ia), mb = param_monomial(ib);
                    const int qq = ma.q + mb.q, d = ma.p + mb.p + qq;
                    sh.H[e] = hsc * sd_half(ia) * sd_half(ib) * sh.mom[ma.t + mb.t][d * (d + 1) / 2 + qq];
                }
            }
            WAVE_SYNC();
            if constexpr (Rules::lm) {
                if (lane == 0) {
                    double mx = sh.H[0];
                    for (int i = 1; i < P; ++i) mx = fmax(mx, sh.H[i * P + i]);
                    const double lambda = 1e-4 * mx;
                    for (int i = 0; i < P; ++i) sh.H[i * P + i] += lambda;
                }
                WAVE_SYNC();
            }
            int ok, pos;
            const bool fast = ldlt_operator_fast<P>(sh, lane, ok, pos);
            if (!fast) { // tied or NaN diagonal: Eigen's exchanges step by step
                if constexpr (P == 12) {
                    ldlt_factor_wave<P>(sh, lane, ok, pos);
                } else {
                    if (lane == 0) {
                        ldlt_factor<P>(sh.H, sh.tr, sh.x, &ok, &pos);
                        sh.state = ok && pos;
                    }
                    WAVE_SYNC();
                    ok = pos = sh.state;
                }
            }
            if (!(ok && pos)) {
                status = DIC_NUMERICAL_FAILURE;
                failed = true;
                break;
            }
            if (!fast) {
                ldlt_operator_column<P>(sh, lane);
                WAVE_SYNC();
            }
            corr = sh.red[P];
            // dp = +H^-1 G (:351-355): lane i forms row i of the operator product
            if (lane < P) {
                double d = 0.0;
#pragma unroll
                for (int j = 0; j < P; ++j) d += sh.Hinv[lane * P + j] * sh.red[j];
                sh.dp[lane] = d;
            }
            WAVE_SYNC();
            // lane 0: cap and compose (:356-390)
            if (lane == 0) {
                double dp[12] = {0.0};
#pragma unroll
                for (int j = 0; j < P; ++j) dp[j] = sh.dp[j];
                int st = 0; // 0 go on, 1 converged, 2 failed
                bool fin = true;
                double nrm = 0.0;
                for (int j = 0; j < P; ++j) {
                    fin = fin && isfinite(dp[j]);
                    nrm += dp[j] * dp[j];
                }
                if (!fin) {
                    st = 2;
                } else {
                    nrm = sqrt(nrm);
                    double test = nrm;
                    if (nrm > Rules::cap) {
                        const double sc = Rules::cap / nrm;
                        for (int j = 0; j < P; ++j) dp[j] *= sc;
                        if (!Rules::precap) test = Rules::cap;
                    }
                    double pn[12] = {0.0};
                    fgn_compose<P>(sh.p, dp, pn);
                    bool okp = true;
                    for (int j = 0; j < P; ++j) {
                        sh.p[j] = pn[j];
                        okp = okp && isfinite(pn[j]);
                    }
                    if (!okp) st = 2;
                    else if (test < a.thr) st = 1;
                }
                sh.state = st;
            }
            WAVE_SYNC();
            const int st = sh.state;
            if (st == 2) {
                status = DIC_NUMERICAL_FAILURE;
                failed = true;
                break;
            }
            if (st == 1) {
                converged = true;
                break;
            }
        }
        if (!failed) {
            if (!converged && isfinite(corr)) converged = true;
            status = converged ? DIC_SUCCESS : DIC_NOT_CONVERGED;
            solved = true;
        }
    }

    if (lane == 0) {
        const dic_initial ini = a.init[poi];
        dic_result r;
        r.x = px;
        r.y = py;
#pragma unroll
        for (int j = 0; j < 12; ++j) r.p[j] = 0.0;
        if (solved) {
#pragma unroll
            for (int j = 0; j < P; ++j) r.p[j] = sh.p[j];
            r.correlation = corr;
        } else {
            r.p[0] = ini.u;
            r.p[1] = ini.v;
            r.correlation = ini.confidence;
        }
        r.status = status;
        r.valid = (status == DIC_SUCCESS) ? 1 : 0;
        r.iterations = iterations;
        r.reserved = 0;
        a.out[poi] = r;
    }
}

template <int P, int OBJ, int D, typename R, bool MASKED>
__global__ __launch_bounds__(64, (kFgnMinWaves<P, R>)) void fgn_kernel(IcgnArgs a)
{
    __shared__ IcgnShared sh;
    const int per = (a.n + 7) >> 3;
    const int poi = (int)(blockIdx.x & 7u) * per + (int)(blockIdx.x >> 3);
    if (poi >= a.n) return;
    fgn_solve_poi<P, OBJ, D, R, MASKED>(a, poi, sh);
}

// Batch FGN descriptors: one POI per workgroup, item after item (the
// dispatch order the item table holds).
template <int P, int OBJ, int D, typename R>
__global__ __launch_bounds__(64, (kFgnMinWaves<P, R>)) void fgn_batch_kernel(const BatchPair *__restrict__ pairs,
                                                                             const int2 *__restrict__ items,
                                                                             int nitems, unsigned *, int)
{
    __shared__ IcgnShared sh;
    if ((int)blockIdx.x >= nitems) return;
    const int2 it = items[blockIdx.x];
    const int pid = __builtin_amdgcn_readfirstlane(it.x);
    const int poi = __builtin_amdgcn_readfirstlane(it.y);
    fgn_solve_poi<P, OBJ, D, R, false>(pairs[pid].a, poi, sh);
}

// Serving ticks with FGN: as icgn_range_kernel, one POI per workgroup.
template <int P, int OBJ, int D, typename R>
__global__ __launch_bounds__(64, (kFgnMinWaves<P, R>)) void fgn_range_kernel(const BatchPair *__restrict__ pairs,
                                                                             const int *__restrict__ offsets,
                                                                             int npairs)
{
    __shared__ IcgnShared sh;
    const int g = (int)blockIdx.x;
    if (g >= offsets[npairs]) return;
    int lo = 0, hi = npairs;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (offsets[mid] <= g) lo = mid;
        else hi = mid;
    }
    const int pid = __builtin_amdgcn_readfirstlane(lo);
    fgn_solve_poi<P, OBJ, D, R, false>(pairs[pid].a, g - offsets[pid], sh);
}

template <int P, int OBJ, int D, typename R, bool MASKED>
hipError_t launch_fgn_variant(const IcgnArgs &args, hipStream_t s)
{
    hipLaunchKernelGGL((fgn_kernel<P, OBJ, D, R, MASKED>), dim3(8 * ((args.n + 7) / 8)), dim3(64), 0, s, args);
    return hipGetLastError();
}

template <int P, int OBJ, int D, typename R, bool MASKED>
hipError_t launch_variant(const IcgnArgs &args, int shape, hipStream_t s)
{
    if (args.fgn) return launch_fgn_variant<P, OBJ, D, R, MASKED>(args, s);
    constexpr bool TWOPASS = OBJ == DIC_OBJECTIVE_ZNSSD && sizeof(R) == 8;
    if (shape == kShapeCoop) {
        const size_t lds = TWOPASS ? sizeof(double) * (size_t)args.nfull : 0;
        auto kern = icgn_coop_kernel<P, OBJ, D, R, MASKED>;
        if (lds > 48 * 1024) {
            hipError_t e =
                hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(kern, dim3(args.n), dim3(64 * kCoopWaves), lds, s, args);
        return hipGetLastError();
    }
    if (P == 6 && sizeof(R) == 4 && !MASKED && shape == kShapeTile) {
        const int groups = (args.n + kGroupPois - 1) / kGroupPois;
        hipLaunchKernelGGL((icgn_group_kernel<P, OBJ, D>), dim3(8 * ((groups + 7) / 8)), dim3(64 * kGroupPois), 0, s,
                           args);
        return hipGetLastError();
    }
    const size_t lds = TWOPASS ? sizeof(double) * (size_t)args.nfull : 0;
    auto kern = icgn_kernel<P, OBJ, D, R, MASKED>;
    if (lds > 48 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kern, dim3(8 * ((args.n + 7) / 8)), dim3(64), lds, s, args);
    return hipGetLastError();
}

template <int P, int OBJ, int D, typename R>
hipError_t dispatch_masked(const IcgnArgs &args, bool masked, int shape, hipStream_t s)
{
    return masked ? launch_variant<P, OBJ, D, R, true>(args, shape, s)
                  : launch_variant<P, OBJ, D, R, false>(args, shape, s);
}

template <int P, int OBJ, int D>
hipError_t dispatch_precision(const IcgnArgs &args, int precision, bool masked, int shape, hipStream_t s)
{
    return precision == DIC_PRECISION_FP64 ? dispatch_masked<P, OBJ, D, double>(args, masked, shape, s)
                                           : dispatch_masked<P, OBJ, D, float>(args, masked, shape, s);
}

template <int P, int OBJ>
hipError_t dispatch_degree(const IcgnArgs &args, int degree, int precision, bool masked, int shape, hipStream_t s)
{
    switch (degree) {
    case 1: return dispatch_precision<P, OBJ, 1>(args, precision, masked, shape, s);
    case 3: return dispatch_precision<P, OBJ, 3>(args, precision, masked, shape, s);
    case 5: return dispatch_precision<P, OBJ, 5>(args, precision, masked, shape, s);
    default: return hipErrorInvalidValue;
    }
}

} // namespace

// static LDS of the solve kernels: their per-POI blocks plus 256 bytes for the
// compiler's alignment of the __shared__ objects
size_t solve_static_lds() { return sizeof(IcgnShared) + 256; }
size_t coop_static_lds() { return sizeof(IcgnShared) + sizeof(CoopSlots<kCoopWaves>) + 256; }

double tile_fit_share(const double *xy, const dic_initial *init, int n, int radius, int degree)
{
    // the box of grow_box over each quad, and stage_tile's size test (before
    // its clamp to the image, which can only shrink the box)
    const int quads = (n + kGroupPois - 1) / kGroupPois, e = radius + radius / 10 + 3, ns = degree + 1;
    if (quads == 0) return 0.0;
    int fit = 0;
    for (int q = 0; q < quads; ++q) {
        int xlo = kBigBox, ylo = kBigBox, xhi = -kBigBox, yhi = -kBigBox;
        for (int k = q * kGroupPois; k < std::min(n, (q + 1) * kGroupPois); ++k) {
            if (!(std::fabs(init[k].u) < 1e6 && std::fabs(init[k].v) < 1e6)) continue;
            if (!(std::fabs(xy[2 * k]) < 1e9 && std::fabs(xy[2 * k + 1]) < 1e9)) continue;
            const int x = (int)std::lround(xy[2 * k]) + (int)std::floor(init[k].u);
            const int y = (int)std::lround(xy[2 * k + 1]) + (int)std::floor(init[k].v);
            xlo = std::min(xlo, x - e);
            xhi = std::max(xhi, x + e);
            ylo = std::min(ylo, y - e);
            yhi = std::max(yhi, y + e);
        }
        if (xhi >= xlo && yhi >= ylo && xhi - xlo + ns <= kGroupPitch && yhi - ylo + ns <= kGroupRows) ++fit;
    }
    return (double)fit / (double)quads;
}

hipError_t launch_icgn(const IcgnLaunch &L, hipStream_t s)
{
    if (L.n <= 0) return hipSuccess;
    const dic_icgn_params &pr = *L.params;
    const int P = pr.shape_order == 2 ? 12 : 6;
    IcgnArgs a;
    a.W = L.width;
    a.H = L.height;
    a.pitch = L.pitch;
    a.b = L.border;
    a.ref = L.ref;
    a.gx = L.gx;
    a.gy = L.gy;
    a.coeffs = L.coeffs;
    a.roi = L.roi;
    a.table = L.table;
    a.nfull = L.table_count;
    a.coeff_bytes = (int)L.coeff_bytes;
    a.radius = pr.subset_radius;
    a.max_it = pr.max_iterations;
    a.thr = pr.convergence_threshold;
    int needed = P;
    if (pr.masked) {
        int ms = (int)std::ceil(pr.min_valid_sample_ratio * (double)L.table_count);
        if (pr.min_valid_samples > ms) ms = pr.min_valid_samples;
        if (ms > needed) needed = ms;
    }
    a.min_needed = needed;
    a.fgn = pr.optimizer == DIC_OPTIMIZER_FGN ? 1 : 0;
    a.xy = L.poi_xy;
    a.init = L.init;
    a.out = L.out;
    a.n = L.n;
    const bool masked = pr.masked != 0;
    const bool ssd = pr.objective == DIC_OBJECTIVE_SSD;
    // first-order FP32: four POIs per workgroup over an LDS tile where the
    // caller's POI layout lets four neighbours share one (tile_hint), else,
    // with no layout known, wherever the radius allows it
    const int shape = L.tile_hint < 0 ? (grouped_radius(pr.subset_radius) ? kShapeTile : kShapeWave) : L.tile_hint;
    if (P == 6)
        return ssd ? dispatch_degree<6, DIC_OBJECTIVE_SSD>(a, L.degree, L.precision, masked, shape, s)
                   : dispatch_degree<6, DIC_OBJECTIVE_ZNSSD>(a, L.degree, L.precision, masked, shape, s);
    return ssd ? dispatch_degree<12, DIC_OBJECTIVE_SSD>(a, L.degree, L.precision, masked, shape, s)
               : dispatch_degree<12, DIC_OBJECTIVE_ZNSSD>(a, L.degree, L.precision, masked, shape, s);
}


namespace {
template <int P, int OBJ, int D, bool PERS>
const void *batch_func_prec(int precision)
{
    return precision == DIC_PRECISION_FP64 ? (const void *)icgn_batch_kernel<P, OBJ, D, double, PERS>
                                           : (const void *)icgn_batch_kernel<P, OBJ, D, float, PERS>;
}
template <int P, int OBJ, bool PERS>
const void *batch_func_deg(int degree, int precision)
{
    switch (degree) {
    case 1: return batch_func_prec<P, OBJ, 1, PERS>(precision);
    case 3: return batch_func_prec<P, OBJ, 3, PERS>(precision);
    case 5: return batch_func_prec<P, OBJ, 5, PERS>(precision);
    default: return nullptr;
    }
}
template <int P, bool PERS>
const void *batch_func_obj(int objective, int degree, int precision)
{
    return objective == DIC_OBJECTIVE_SSD ? batch_func_deg<P, DIC_OBJECTIVE_SSD, PERS>(degree, precision)
                                          : batch_func_deg<P, DIC_OBJECTIVE_ZNSSD, PERS>(degree, precision);
}
const void *batch_func(int order, bool persistent, int objective, int degree, int precision)
{
    if (order == 2)
        return persistent ? batch_func_obj<12, true>(objective, degree, precision)
                          : batch_func_obj<12, false>(objective, degree, precision);
    return persistent ? batch_func_obj<6, true>(objective, degree, precision)
                      : batch_func_obj<6, false>(objective, degree, precision);
}

template <int P, int OBJ, int D>
const void *fgn_batch_func_prec(int precision)
{
    return precision == DIC_PRECISION_FP64 ? (const void *)fgn_batch_kernel<P, OBJ, D, double>
                                           : (const void *)fgn_batch_kernel<P, OBJ, D, float>;
}
template <int P, int OBJ>
const void *fgn_batch_func_deg(int degree, int precision)
{
    switch (degree) {
    case 1: return fgn_batch_func_prec<P, OBJ, 1>(precision);
    case 3: return fgn_batch_func_prec<P, OBJ, 3>(precision);
    case 5: return fgn_batch_func_prec<P, OBJ, 5>(precision);
    default: return nullptr;
    }
}
const void *fgn_batch_func(int order, int objective, int degree, int precision)
{
    const bool ssd = objective == DIC_OBJECTIVE_SSD;
    if (order == 2)
        return ssd ? fgn_batch_func_deg<12, DIC_OBJECTIVE_SSD>(degree, precision)
                   : fgn_batch_func_deg<12, DIC_OBJECTIVE_ZNSSD>(degree, precision);
    return ssd ? fgn_batch_func_deg<6, DIC_OBJECTIVE_SSD>(degree, precision)
               : fgn_batch_func_deg<6, DIC_OBJECTIVE_ZNSSD>(degree, precision);
}

template <int P, int OBJ, int D>
const void *fgn_range_func_prec(int precision)
{
    return precision == DIC_PRECISION_FP64 ? (const void *)fgn_range_kernel<P, OBJ, D, double>
                                           : (const void *)fgn_range_kernel<P, OBJ, D, float>;
}
template <int P, int OBJ>
const void *fgn_range_func_deg(int degree, int precision)
{
    switch (degree) {
    case 1: return fgn_range_func_prec<P, OBJ, 1>(precision);
    case 3: return fgn_range_func_prec<P, OBJ, 3>(precision);
    case 5: return fgn_range_func_prec<P, OBJ, 5>(precision);
    default: return nullptr;
    }
}
const void *fgn_range_func(int order, int objective, int degree, int precision)
{
    const bool ssd = objective == DIC_OBJECTIVE_SSD;
    if (order == 2)
        return ssd ? fgn_range_func_deg<12, DIC_OBJECTIVE_SSD>(degree, precision)
                   : fgn_range_func_deg<12, DIC_OBJECTIVE_ZNSSD>(degree, precision);
    return ssd ? fgn_range_func_deg<6, DIC_OBJECTIVE_SSD>(degree, precision)
               : fgn_range_func_deg<6, DIC_OBJECTIVE_ZNSSD>(degree, precision);
}

template <int OBJ>
const void *batch_group_func_deg(int degree)
{
    switch (degree) {
    case 1: return (const void *)icgn_batch_group_kernel<OBJ, 1>;
    case 3: return (const void *)icgn_batch_group_kernel<OBJ, 3>;
    case 5: return (const void *)icgn_batch_group_kernel<OBJ, 5>;
    default: return nullptr;
    }
}

struct Launch1 { // one group's launch
    const void *func;
    size_t lds;
    const int2 *items;
    int nitems;
    unsigned *counter;
    unsigned grid, block;
};

hipError_t batch_launch_of(const BatchLaunch &L, int g, Launch1 &o)
{
    const BatchGroup &G = L.groups[g];
    const bool twopass = L.objective == DIC_OBJECTIVE_ZNSSD && L.precision == DIC_PRECISION_FP64;
    o.items = L.items + G.first;
    o.nitems = G.count;
    o.counter = L.counter + g;
    if (G.fgn) { // one POI per workgroup whatever the mapping
        o.func = fgn_batch_func(G.shape_order, L.objective, L.degree, L.precision);
        if (!o.func) return hipErrorInvalidValue;
        o.lds = solve_dynamic_lds(true, L.precision, L.objective, G.max_nfull);
        o.grid = (unsigned)G.count;
        o.block = 64u;
        if (o.lds > 48 * 1024)
            return hipFuncSetAttribute(o.func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)o.lds);
        return hipSuccess;
    }
    if (G.tile) {
        o.func = L.objective == DIC_OBJECTIVE_SSD ? batch_group_func_deg<DIC_OBJECTIVE_SSD>(L.degree)
                                                  : batch_group_func_deg<DIC_OBJECTIVE_ZNSSD>(L.degree);
        if (!o.func) return hipErrorInvalidValue;
        o.lds = 0;
        const unsigned quads = (unsigned)((G.count + kGroupPois - 1) / kGroupPois);
        o.grid = 8u * ((quads + 7u) / 8u);
        o.block = 64u * kGroupPois;
        return hipSuccess;
    }
    o.func = batch_func(G.shape_order, L.persistent, L.objective, L.degree, L.precision);
    if (!o.func) return hipErrorInvalidValue;
    o.lds = twopass ? sizeof(double) * (size_t)G.max_nfull : 0;
    o.grid = L.persistent ? (unsigned)L.grid : (unsigned)G.count;
    o.block = 64u;
    if (o.lds > 48 * 1024) return hipFuncSetAttribute(o.func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)o.lds);
    return hipSuccess;
}
} // namespace

hipError_t launch_batch(const BatchLaunch &L, hipStream_t s)
{
    hipError_t e;
    if (L.persistent) {
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, s, L.counter, L.ngroups, 0u);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    for (int g = 0; g < L.ngroups; ++g) {
        Launch1 l;
        e = batch_launch_of(L, g, l);
        if (e != hipSuccess) return e;
        if (l.nitems <= 0) continue;
        const BatchPair *pairs = L.pairs;
        const int2 *items = l.items;
        int nitems = l.nitems, chunk = L.chunk;
        unsigned *counter = l.counter;
        void *args[] = {(void *)&pairs, (void *)&items, (void *)&nitems, (void *)&counter, (void *)&chunk};
        e = hipLaunchKernel(l.func, dim3(l.grid), dim3(l.block), args, l.lds, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t add_batch_nodes(const BatchLaunch &L, hipGraph_t g)
{
    hipError_t e;
    hipGraphNode_t prev = nullptr;
    if (L.persistent) { // the queue counters are cleared by a kernel node, not a memset node
        unsigned *counter = L.counter;
        int n = L.ngroups;
        unsigned zero = 0;
        void *rargs[] = {(void *)&counter, (void *)&n, (void *)&zero};
        hipKernelNodeParams rp{};
        rp.func = (void *)fill_u32_kernel;
        rp.gridDim = dim3(1);
        rp.blockDim = dim3(64);
        rp.sharedMemBytes = 0;
        rp.kernelParams = rargs;
        rp.extra = nullptr;
        e = hipGraphAddKernelNode(&prev, g, nullptr, 0, &rp);
        if (e != hipSuccess) return e;
    }
    for (int k = 0; k < L.ngroups; ++k) { // groups in the host's order (heavier first)
        Launch1 l;
        e = batch_launch_of(L, k, l);
        if (e != hipSuccess) return e;
        if (l.nitems <= 0) continue;
        const BatchPair *pairs = L.pairs;
        const int2 *items = l.items;
        int nitems = l.nitems, chunk = L.chunk;
        unsigned *counter = l.counter;
        void *kargs[] = {(void *)&pairs, (void *)&items, (void *)&nitems, (void *)&counter, (void *)&chunk};
        hipKernelNodeParams kp{};
        kp.func = const_cast<void *>(l.func);
        kp.gridDim = dim3(l.grid);
        kp.blockDim = dim3(l.block);
        kp.sharedMemBytes = (unsigned)l.lds;
        kp.kernelParams = kargs;
        kp.extra = nullptr;
        hipGraphNode_t node;
        e = hipGraphAddKernelNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0, &kp);
        if (e != hipSuccess) return e;
        prev = node;
    }
    return hipSuccess;
}


namespace {
template <int P, int OBJ, int D>
const void *range_func_prec(int precision)
{
    if constexpr (P == 6)
        if (precision != DIC_PRECISION_FP64) return (const void *)icgn_range_group_kernel<OBJ, D>;
    return precision == DIC_PRECISION_FP64 ? (const void *)icgn_range_kernel<P, OBJ, D, double>
                                           : (const void *)icgn_range_kernel<P, OBJ, D, float>;
}
template <int P, int OBJ>
const void *range_func_deg(int degree, int precision)
{
    switch (degree) {
    case 1: return range_func_prec<P, OBJ, 1>(precision);
    case 3: return range_func_prec<P, OBJ, 3>(precision);
    case 5: return range_func_prec<P, OBJ, 5>(precision);
    default: return nullptr;
    }
}
template <int P>
const void *range_func_obj(int objective, int degree, int precision)
{
    return objective == DIC_OBJECTIVE_SSD ? range_func_deg<P, DIC_OBJECTIVE_SSD>(degree, precision)
                                          : range_func_deg<P, DIC_OBJECTIVE_ZNSSD>(degree, precision);
}
} // namespace

hipError_t launch_range(const RangeLaunch &L, hipStream_t s)
{
    if (L.total <= 0) return hipSuccess;
    if (L.fgn) {
        const void *ff = fgn_range_func(L.shape_order, L.objective, L.degree, L.precision);
        if (!ff) return hipErrorInvalidValue;
        const size_t fl = solve_dynamic_lds(true, L.precision, L.objective, L.max_nfull);
        if (fl > 48 * 1024) {
            hipError_t e = hipFuncSetAttribute(ff, hipFuncAttributeMaxDynamicSharedMemorySize, (int)fl);
            if (e != hipSuccess) return e;
        }
        const BatchPair *pairs = L.pairs;
        const int *offsets = L.offsets;
        int npairs = L.npairs;
        void *args[] = {(void *)&pairs, (void *)&offsets, (void *)&npairs};
        return hipLaunchKernel(ff, dim3((unsigned)L.total), dim3(64), args, fl, s);
    }
    const void *func = L.shape_order == 2 ? range_func_obj<12>(L.objective, L.degree, L.precision)
                                          : range_func_obj<6>(L.objective, L.degree, L.precision);
    if (!func) return hipErrorInvalidValue;
    const bool twopass = L.objective == DIC_OBJECTIVE_ZNSSD && L.precision == DIC_PRECISION_FP64;
    const size_t lds = twopass ? sizeof(double) * (size_t)L.max_nfull : 0;
    if (lds > 48 * 1024) {
        hipError_t e = hipFuncSetAttribute(func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    const BatchPair *pairs = L.pairs;
    const int *offsets = L.offsets;
    int npairs = L.npairs;
    void *args[] = {(void *)&pairs, (void *)&offsets, (void *)&npairs};
    if (L.shape_order == 1 && L.precision != DIC_PRECISION_FP64) // icgn_range_group_kernel
        return hipLaunchKernel(func, dim3((unsigned)((L.total + kGroupPois - 1) / kGroupPois)), dim3(64 * kGroupPois),
                               args, 0, s);
    return hipLaunchKernel(func, dim3((unsigned)L.total), dim3(64), args, lds, s);
}

hipError_t launch_init_from_results(const dic_result *prev, dic_initial *init, int n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(init_from_results_kernel, dim3((n + 255) / 256), dim3(256), 0, s, prev, init, n);
    return hipGetLastError();
}

hipError_t launch_init_extrapolated(const dic_result *a, const dic_result *b, dic_initial *init, int n,
                                    hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(init_extrapolated_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a, b, init, n);
    return hipGetLastError();
}

} // namespace dic
