// Wave-autonomous moment fold (precision form, blackbody atmosphere).
//
// Reference behaviour: the energy x cell x phase spline loop of
// xpsi/cellmesh/integrator_for_azimuthal_invariance.pyx:565-596 with the
// blackbody intensity of xpsi/surface_radiation_field/hot_BB.pyx:54-98, in
// the moment form of k_moment_fold (xh_moment.hip): where the periodic Akima
// piece s_k of energy k is >= 0 over its whole interval i,
//
//     sum_j A_j s_k(x_jp) = sum_i sum_m coef_k[i][m] * MOM[i][m][p],
//     MOM[i][m][p]       = sum over cells j with i(j, p) = i of A_j t^m,
//
// one 16x16x4 f32 MFMA per interval, 16 energies and 16 phases.  Pieces that
// change sign are evaluated point by point with the clipping; pieces <= 0
// throughout fold nothing.
//
// Here one wavefront owns one member, 64 energies, up to 64 phases and a
// block of parallels, and needs no other wave: per parallel and image it
// walks the intervals once, in order, and for each interval
//   - lanes = energies: the profile at the next knot, the next secant, the
//     Akima coefficients of the piece from a sliding window of five secants
//     and its sign class, all in registers;
//   - lanes = phases: the cells whose phase point falls in the interval, in
//     azimuth order from the lane's own rotation point, summed into the four
//     moments, also in registers;
//   - both sets of four registers transposed by row blocks with the gfx950
//     lane swaps (v_permlane32_swap, v_permlane16_swap) into the MFMA's A
//     (16 energies x 4 powers, one per 16-energy chunk) and B (4 powers x 16
//     phases, one per 16-phase tile) operands, and the 4 x 4 products issued;
//   - sign-changing pieces: the coefficients of each such energy read from
//     its lane, and every phase lane evaluates its own cells of the interval.
// There are no block barriers and no LDS tables of coefficients or moments:
// LDS holds only the parallel's cell azimuths and areas and the image's leaf
// geometry.  The FP32 sums of each accumulation chain (a parallel's
// intervals, cut every F.chain MFMA steps so that the rounding of a long chain
// does not add up into an error shared by neighbouring bins) are scaled in
// FP64 and added into the flux with FP64 atomics, so parallels of one member
// may be spread over any number of waves; the flux is zeroed before the
// launch.
//
// Cell azimuths must increase along a parallel and span less than 2 pi (the
// equal-area meshes do): then each phase lane's wrapped points are one
// rotation of a sorted list.  A parallel that breaks this marks the member's
// status (a loud failure, never a silent wrong sum).
#include "xh_common.hpp"
#include "xh_types.hpp"
#include "xh_kernels.hpp"

namespace xh {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

// The 4 x 4 transpose of (register, 16-lane row) blocks: on return row r of
// v[c] holds what row c of v[r] held.
__device__ __forceinline__ void row_transpose(float (&v)[4])
{
    unsigned a = __float_as_uint(v[0]), b = __float_as_uint(v[1]), c = __float_as_uint(v[2]),
             d = __float_as_uint(v[3]);
    auto ac = __builtin_amdgcn_permlane32_swap(a, c, false, false);
    auto bd = __builtin_amdgcn_permlane32_swap(b, d, false, false);
    auto ab = __builtin_amdgcn_permlane16_swap(ac[0], bd[0], false, false);
    auto cd = __builtin_amdgcn_permlane16_swap(ac[1], bd[1], false, false);
    v[0] = __uint_as_float(ab[0]);
    v[1] = __uint_as_float(ab[1]);
    v[2] = __uint_as_float(cd[0]);
    v[3] = __uint_as_float(cd[1]);
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ float readlane_f(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ double readlane_d(double v, int l)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// sign class of a + t (b + t (c + t d)) on [0, h], as k_moment_fold's
// cubic_class<float> (1: >= -tol throughout, 0: <= tol throughout, 2: mixed)
__device__ __forceinline__ int piece_class(float a, float b, float c, float d, float h)
{
    if (a == 0.f && b == 0.f && c == 0.f && d == 0.f) return 0;
    const float bh = b * h, ch = c * h * h, dh = d * h * h * h;
    const float scale = fabsf(a) + fabsf(bh) + fabsf(ch) + fabsf(dh);
    const float tol = 64.f * FLT_EPSILON * scale;
    constexpr float third = 1.f / 3.f;
    const float b0 = a, b1 = a + bh * third, b2 = a + 2.f * bh * third + ch * third, b3 = a + bh + ch + dh;
    float lo = fminf(fminf(b0, b1), fminf(b2, b3)), hi = fmaxf(fmaxf(b0, b1), fmaxf(b2, b3));
    if (lo >= -tol) return 1;
    if (hi <= tol) return 0;
    auto f = [&](float t) { return a + t * (b + t * (c + t * d)); };
    lo = fminf(a, b3);
    hi = fmaxf(a, b3);
    float r1 = -1.f, r2 = -1.f;
    if (d != 0.f) {
        const float disc = c * c - 3.f * b * d;
        if (disc >= 0.f) {
            const float q = -(c + copysignf(sqrtf(disc), c));
            if (q != 0.f) { r1 = q / (3.f * d); r2 = b / q; }
            else r1 = 0.f;
        }
    } else if (c != 0.f) r1 = -b / (2.f * c);
    if (r1 > 0.f && r1 < h) { const float v = f(r1); lo = fminf(lo, v); hi = fmaxf(hi, v); }
    if (r2 > 0.f && r2 < h) { const float v = f(r2); lo = fminf(lo, v); hi = fmaxf(hi, v); }
    if (lo >= -tol) return 1;
    if (hi <= tol) return 0;
    return 2;
}

}  // namespace

// grid: x = ring block, y = energy chunk x phase chunk (64 each), z = member;
// 64 threads; dynamic LDS wave_fold_lds(N_L, S_max)
__global__ void __launch_bounds__(64, 4) k_wave_fold(Member* members, Layout L, FoldArgs F)
{
    const int mi = blockIdx.z, lane = threadIdx.x;
    if (!members[mi].active || members[mi].status) return;
    const int S = members[mi].S, correct = members[mi].correct;
    const double else_logT = members[mi].else_logT;
    const int N_L = L.N_L, N_P = L.N_P, N_E = L.N_E, NI = N_L - 1;
    const int n_pc = (N_P + 63) >> 6;
    const int e0 = (blockIdx.y / n_pc) * 64, p0 = (blockIdx.y % n_pc) * 64;
    const int r0 = blockIdx.x * F.ring_block;
    if (r0 >= S || e0 >= N_E) return;
    const int r1 = min(S, r0 + F.ring_block);

    extern __shared__ __align__(16) double wsm[];
    double2* KN = (double2*)wsm;             // [N_L] knot: (geometric factor, 0 where hidden; 1 / redshift)
    double* PH = (double*)(KN + N_L);        // [N_L] leaf phases
    double* CPHI = PH + N_L;                 // [S_max] cell azimuths
    float2* KH = (float2*)(CPHI + L.S_max);  // [N_L] interval (width, 1 / width) in FP32
    float* CAREA = (float*)(KH + N_L);       // [S_max] area / largest area, NaN = not radiating
    unsigned char* LIVE = (unsigned char*)(CAREA + L.S_max);   // [N_L] piece's stencil shows a leaf

    // energy lane
    const int e = e0 + lane;
    const bool e_ok = e < N_E;
    const double E = F.energies[min(e, N_E - 1)];
    const double esc = BB_NORM / (E * KEV);  // the flux scale of this lane's energy
    const double inv_kT_else = 1.0 / (K_B_OVER_KEV * pow(10.0, else_logT));
    // phase lane
    const int p = p0 + lane;
    const bool p_ok = p < N_P;
    const double ph = p_ok ? F.phases[p] : 0.0;

    v4f acc[4][4];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int t = 0; t < 4; t++) acc[c][t] = v4f{0.f, 0.f, 0.f, 0.f};
    double* flux = F.flux + (size_t)mi * N_E * N_P;
    const size_t cbase = (size_t)mi * L.S_max * L.S_max;

    for (int i = r0; i < r1; i++) {
        const size_t ro = (size_t)mi * L.S_max + i;
        if (!F.valid[ro * L.IO_max]) continue;
        __syncthreads();   // the previous parallel's reads of the cell arrays are done
        double amax = 0.0;
        for (int j = lane; j < S; j += 64) {
            const size_t c = cbase + (size_t)i * S + j;
            const double a = F.cellArea[c];
            const bool rad = F.radiates ? (F.radiates[c] == 1) : (a > 0.0);
            CPHI[j] = F.phi[(size_t)mi * L.S_max * (F.phi_row_stride ? L.S_max : 1) + (size_t)i * F.phi_row_stride + j];
            if (rad && a > 0.0) amax = fmax(amax, a);
        }
        amax = wave_max(amax);
        const double a_ref = amax > 0.0 ? amax : 1.0;
        for (int j = lane; j < S; j += 64) {
            const size_t c = cbase + (size_t)i * S + j;
            const double a = F.cellArea[c];
            const bool rad = F.radiates ? (F.radiates[c] == 1) : (a > 0.0);
            CAREA[j] = rad ? (float)(a / a_ref) : __builtin_nanf("");
        }
        const double inv_kT = 1.0 / (K_B_OVER_KEV * pow(10.0, F.logT[ro]));
        bool any = false;
        int chain_n = 0;   // MFMA steps in the current FP32 chain
        // the chain's sums, scaled in FP64 (a_ref, the blackbody norm,
        // 1 / (E keV)), into the flux; the chain restarts from zero
        auto flush = [&]() {
            // the lane index and the flux row length re-read here, so that the
            // flush's addresses are formed at each flush and not held in
            // registers across the sweep
            int ln = lane, np = N_P;
            asm volatile("" : "+v"(ln));
            asm volatile("" : "+s"(np));
            if (e0 + 64 <= N_E && p0 + 64 <= np) {
                // a full tile: each lane's 64 targets sit at one byte offset
                // from 64 wave-uniform rows, with no bounds to test
                const unsigned boff = (unsigned)((4 * (ln >> 4) * np + (ln & 15)) * 8);
                double* tile = flux + (size_t)e0 * np + p0;
#pragma unroll
                for (int c = 0; c < 4; c++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const double sc = a_ref * __shfl(esc, 16 * c + 4 * (ln >> 4) + r, 64);
#pragma unroll
                        for (int t = 0; t < 4; t++) {
                            double* row = tile + (size_t)((16 * c + r) * np + 16 * t);
                            const float v = acc[c][t][r];
                            // most of a short chain's sums are zero: no atomic for those
                            if (v != 0.f) unsafeAtomicAdd((double*)((char*)row + boff), (double)v * sc);
                        }
                    }
            } else {
#pragma unroll
                for (int c = 0; c < 4; c++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int src = 16 * c + 4 * (ln >> 4) + r, ee = e0 + src;
                        const double sc = a_ref * __shfl(esc, src, 64);
                        if (ee >= N_E) continue;
#pragma unroll
                        for (int t = 0; t < 4; t++) {
                            const int pq = p0 + 16 * t + (ln & 15);
                            const float v = acc[c][t][r];
                            if (pq < np && v != 0.f) unsafeAtomicAdd(flux + (size_t)ee * np + pq, (double)v * sc);
                        }
                    }
            }
#pragma unroll
            for (int c = 0; c < 4; c++)
#pragma unroll
                for (int t = 0; t < 4; t++) acc[c][t] = v4f{0.f, 0.f, 0.f, 0.f};
            any = false;
            chain_n = 0;
        };
        for (int I = 0; I < L.IO_max; I++) {
            const size_t vix = ro * L.IO_max + I;
            if (!F.valid[vix]) break;
            __syncthreads();   // the previous image's reads of the geometry are done
            const double* g = F.geo + vix * 5 * N_L;
            // the last leaf repeats the first when it was not queried itself
            const bool copy_last = g[5 * N_L - 1] < -1.0;
            for (int q = lane; q < N_L; q += 64) {
                const int mm = (q == N_L - 1 && copy_last) ? 0 : q;
                const double geom = g[3 * N_L + mm];
                const bool vis = geom != 0.0 && g[4 * N_L + mm] >= -1.0;
                KN[q] = make_double2(vis ? geom : 0.0, 1.0 / g[N_L + mm]);
                PH[q] = g[q];
                if (q < NI) {
                    const double w = g[q + 1] - g[q];
                    KH[q] = make_float2((float)w, (float)(1.0 / w));
                }
            }
            __syncthreads();
            // pieces whose Akima stencil (secants i-2 .. i+2, periodic) touches a
            // visible leaf; the others are identically zero
            for (int q = lane; q < NI; q += 64) {
                bool any_vis = false;
                for (int k = q - 2; k <= q + 2; k++) {
                    const int sk = k < 0 ? k + NI : (k >= NI ? k - NI : k);
                    any_vis = any_vis || KN[sk].x != 0.0 || KN[sk + 1].x != 0.0;
                }
                LIVE[q] = any_vis ? 1 : 0;
            }
            __syncthreads();
            const double lo = PH[0], hi = PH[N_L - 1];

            // phase lane: the rotation point j0 (first cell after the wrap drop)
            int j0 = 0;
            if (p_ok && S > 0) {
                const double x0 = wrap_phase(ph + CPHI[0], lo, hi);
                int a = 1, b = S;
                while (a < b) {
                    const int mid = (a + b) >> 1;
                    if (wrap_phase(ph + CPHI[mid], lo, hi) < x0) b = mid;
                    else a = mid + 1;
                }
                j0 = a == S ? 0 : a;
            }
            auto cell = [&](int k) { int j = j0 + k; return j >= S ? j - S : j; };
            // Within each of the two runs of the rotation (j >= j0, then j < j0)
            // the wrap adds the same multiple of 2 pi; for one turn (the usual
            // case) x = (ph + phi_j) -/+ 2 pi is then formed exactly as
            // wrap_phase forms it.  A run that needs more turns is evaluated
            // by wrap_phase cell by cell.
            auto turns = [&](int j) {
                const double s0 = ph + CPHI[j], x = wrap_phase(s0, lo, hi);
                return x == s0 ? 0 : (x == s0 - TWO_PI ? -1 : (x == s0 + TWO_PI ? 1 : 2));
            };
            int tA = 0, tB = 0;
            if (p_ok && S > 0) { tA = turns(j0); tB = j0 > 0 ? turns(0) : 0; }
            const double offA = tA == -1 ? -TWO_PI : (tA == 1 ? TWO_PI : 0.0);
            const double offB = tB == -1 ? -TWO_PI : (tB == 1 ? TWO_PI : 0.0);
            const bool slow = tA == 2 || tB == 2;
            auto xof = [&](int j) -> double {
                const double s0 = ph + CPHI[j];
                double x = s0 + (j >= j0 ? offA : offB);   // s0 - 2 pi == s0 + (-2 pi) exactly
                if (slow) x = wrap_phase(s0, lo, hi);
                return x;
            };
            // the next two cells in rotation order, read ahead
            int cur = 0;   // cells consumed
            double xn = __builtin_inf(), xn2 = __builtin_inf();
            float an = 0.f, an2 = 0.f;
            if (p_ok && S > 0) { xn = xof(j0); an = CAREA[j0]; }
            if (p_ok && S > 1) { const int j = cell(1); xn2 = xof(j); an2 = CAREA[j]; }
            bool unsorted = false;

            // energy lane: profile at knot k (intensity x geometry, 0 where hidden)
            auto knot = [&](double2 kn) -> float {
                if (kn.x == 0.0 || !e_ok) return 0.f;
                const double Ep = E * kn.y;
                const float Ef = (float)Ep;
                const float Ep3 = Ef * Ef * Ef;
                // IEEE division: a one-signed rounding here would scale every bin alike
                float I_E = Ep3 / (expf((float)(Ep * inv_kT)) - 1.0f);
                if (correct) I_E -= Ep3 / (expf((float)(Ep * inv_kT_else)) - 1.0f);
                return I_E * (float)kn.x;
            };
            // window: y_i, y_{i+1}, y_{i+2} and secants m_{i-2} .. m_{i+1}
            const float yl2 = knot(KN[NI - 2]), yl1 = knot(KN[NI - 1]), yl0 = knot(KN[NI]);
            float ya = knot(KN[0]), yb = knot(KN[1]), yk = knot(KN[2]);
            // secants and the cubic's coefficients by the interval's reciprocal
            // width, rounded once from FP64: a fixed factor per interval whose
            // sign varies from interval to interval (no bias shared by the bins)
            float sm2 = (yl1 - yl2) * KH[NI - 2].y, sm1 = (yl0 - yl1) * KH[NI - 1].y;
            float s0 = (yb - ya) * KH[0].y, s1 = (yk - yb) * KH[1].y;
            const float s0_first = s0, s1_first = s1;
            // Akima weights carried from piece to piece: w of piece i is wn of
            // piece i - 1, and its al that piece's aln (recomputed after a
            // skipped piece)
            float w_c = 0.f, al_c = 0.f;
            bool carry = false;
            double2 kn_next = KN[min(3, NI)];

            for (int iv = 0; iv < NI; iv++) {
                // -- lanes = energies: the next knot and secant (the window slides always)
                const int k2 = iv + 2;
                const double2 kn_now = kn_next;
                if (k2 + 2 <= NI) kn_next = KN[k2 + 2];   // read a step ahead
                float s2, ynew = 0.f;
                if (k2 < NI) {
                    ynew = knot(kn_now);
                    s2 = (ynew - yk) * KH[k2].y;
                } else s2 = k2 == NI ? s0_first : s1_first;
                const bool piece_live = LIVE[iv] != 0;

                // -- lanes = phases: this interval's cells (their moments if the piece can fold)
                const double lb = PH[iv], ub = iv < NI - 1 ? PH[iv + 1] : __builtin_inf();
                float M[4] = {0.f, 0.f, 0.f, 0.f};
                const int c_first = cur;
                int cnt = 0;
                while (xn < ub) {
                    unsorted = unsorted || xn < lb || xn > hi;
                    if (piece_live && !isnan(an)) {
                        const float tq = (float)(xn - lb);
                        float at = an;
                        M[0] += at; at *= tq; M[1] += at; at *= tq; M[2] += at; at *= tq; M[3] += at;
                        cnt++;
                    }
                    cur++;
                    xn = xn2;
                    an = an2;
                    if (cur + 1 < S) { const int j = cell(cur + 1); xn2 = xof(j); an2 = CAREA[j]; }
                    else xn2 = __builtin_inf();
                }
                const int c_last = cur;
                const unsigned long long nzB = __ballot(cnt > 0);

                const float wn = fabsf(s2 - s1) + fabsf(s0 - sm1);
                if (piece_live && nzB) {
                    // -- lanes = energies: the piece's coefficients and class
                    if (!carry) {
                        w_c = fabsf(s1 - s0) + fabsf(sm1 - sm2);
                        al_c = w_c != 0.f ? fabsf(sm1 - sm2) / w_c : 0.f;
                    }
                    const float h = KH[iv].x, ih = KH[iv].y;
                    const float ca = ya;
                    float cb, cc, cd;
                    const float aln = wn != 0.f ? fabsf(s0 - sm1) / wn : 0.f;
                    if (w_c == 0.f) { cb = s0; cc = 0.f; cd = 0.f; }
                    else {
                        const float tn = wn == 0.f ? s0 : (1.f - aln) * s0 + aln * s1;
                        cb = (1.f - al_c) * sm1 + al_c * s0;
                        cc = (3.f * s0 - 2.f * cb - tn) * ih;
                        cd = (cb + tn - 2.f * s0) * (ih * ih);
                    }
                    w_c = wn;
                    al_c = aln;
                    carry = true;
                    const int cls = correct ? 1 : piece_class(ca, cb, cc, cd, h);
                    float P[4];
                    const bool on = cls == 1 && e_ok;
                    P[0] = on ? ca : 0.f; P[1] = on ? cb : 0.f; P[2] = on ? cc : 0.f; P[3] = on ? cd : 0.f;
                    const unsigned long long nzA =
                        __ballot(on && (ca != 0.f || cb != 0.f || cc != 0.f || cd != 0.f));
                    const unsigned long long mixA = __ballot(cls == 2 && e_ok);

                    if (nzA) {
                        row_transpose(P);   // P[c]: A operand of energy chunk c
                        row_transpose(M);   // M[t]: B operand of phase tile t
#pragma unroll
                        for (int c = 0; c < 4; c++)
#pragma unroll
                            for (int t = 0; t < 4; t++)
                                acc[c][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(P[c], M[t], acc[c][t], 0, 0, 0);
                        any = true;
                        // a bounded chain: its rounding stays local to its bins
                        if (++chain_n == F.chain) flush();
                    }

                    // -- sign-changing pieces, point by point with the clipping
                    if (mixA) {
                        unsigned long long mb = mixA;
                        while (mb) {
                            const int el = __ffsll((long long)mb) - 1;
                            mb &= mb - 1ull;
                            const float a = readlane_f(ca, el), b = readlane_f(cb, el), c = readlane_f(cc, el),
                                        d = readlane_f(cd, el);
                            const double sc = a_ref * readlane_d(esc, el);
                            float sum = 0.f;
                            if (cnt > 0) {
                                for (int k = c_first; k < c_last; k++) {
                                    const int j = cell(k);
                                    const float ar = CAREA[j];
                                    if (isnan(ar)) continue;
                                    const float tq = (float)(xof(j) - lb);
                                    const float v = a + tq * (b + tq * (c + tq * d));
                                    if (v > 0.f) sum += ar * v;
                                }
                            }
                            if (sum != 0.f && p_ok) unsafeAtomicAdd(flux + (size_t)(e0 + el) * N_P + p, (double)sum * sc);
                        }
                    }
                } else carry = false;
                // slide the window
                ya = yb; yb = yk; yk = ynew;
                sm2 = sm1; sm1 = s0; s0 = s1; s1 = s2;
            }
            // a point below its interval: the azimuths were not one sorted turn
            if (__ballot(unsorted) && lane == 0) members[mi].status = 2;
        }
        // the parallel's last chain
        if (any) flush();
    }
}


// The wave fold with a parallel's cell walk shared by its energy chunks.
// One block per (parallel, 64-phase chunk, member) of NCH = ceil(N_E / 64)
// wavefronts; wavefront w folds energies 64 w .. 64 w + 63.  The intervals
// are taken in slabs of WF_SLAB (16): the block's wavefronts split a slab's intervals,
// each walking the phase lanes' cells through its share (starting from a
// binary search of the rotation order) and writing the lanes' four moments
// per interval to LDS (already in the MFMA B operand's order), the lanes
// with cells as a mask, and each lane's first cell per interval; after a
// barrier every wavefront runs the slab's intervals for its own energies
// exactly as k_wave_fold does (profile, Akima coefficients and class, MFMA
// chain with the same flushes, sign-changing pieces point by point).  The
// walk is done once per parallel instead of once per energy chunk.
__global__ void __launch_bounds__(256, 4) k_wave_fold_shared(Member* members, Layout L, FoldArgs F)
{
    const int mi = blockIdx.z, tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nch = blockDim.x >> 6;
    // the member's status is read once for the whole block: another block of
    // the same member may set it meanwhile, and the block's wavefronts must
    // agree on whether to run (they meet at barriers)
    __shared__ int go_s;
    if (tid == 0) go_s = members[mi].active && !members[mi].status;
    __syncthreads();
    if (!go_s) return;                                       // the whole block
    const int S = members[mi].S, correct = members[mi].correct;
    const int i = blockIdx.x;                                // the parallel
    if (i >= S) return;                                      // the whole block
    const double else_logT = members[mi].else_logT;
    const int N_L = L.N_L, N_P = L.N_P, N_E = L.N_E, NI = N_L - 1;
    const int p0 = blockIdx.y * 64, e0 = w * 64;
    const size_t ro = (size_t)mi * L.S_max + i;
    if (!F.valid[ro * L.IO_max]) return;                     // the whole block

    extern __shared__ __align__(16) double wsm[];
    double2* KN = (double2*)wsm;
    double* PH = (double*)(KN + N_L);
    double* CPHI = PH + N_L;
    float2* KH = (float2*)(CPHI + L.S_max);
    float* CAREA = (float*)(KH + N_L);
    unsigned char* LIVE = (unsigned char*)(CAREA + L.S_max);
    char* tail = (char*)wsm + wave_fold_lds(N_L, L.S_max);
    float* MOM = (float*)tail;                                            // [slab][power][64 phases]
    unsigned long long* MASK = (unsigned long long*)(MOM + WF_SLAB * 4 * 64);   // [slab]
    int* CF = (int*)(MASK + WF_SLAB);                                     // [slab + 1][64]: first cell
    __shared__ double amax_w[4];

    const int e = e0 + lane;
    const bool e_ok = e < N_E;
    const double E = F.energies[min(e, N_E - 1)];
    const double esc = BB_NORM / (E * KEV);
    const double inv_kT_else = 1.0 / (K_B_OVER_KEV * pow(10.0, else_logT));
    const int p = p0 + lane;
    const bool p_ok = p < N_P;
    const double ph = p_ok ? F.phases[p] : 0.0;

    v4f acc[4][4];
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int t = 0; t < 4; t++) acc[c][t] = v4f{0.f, 0.f, 0.f, 0.f};
    double* flux = F.flux + (size_t)mi * N_E * N_P;
    const size_t cbase = (size_t)mi * L.S_max * L.S_max;

    // the parallel's cells, by the block
    double amax = 0.0;
    for (int j = tid; j < S; j += blockDim.x) {
        const size_t c = cbase + (size_t)i * S + j;
        const double a = F.cellArea[c];
        const bool rad = F.radiates ? (F.radiates[c] == 1) : (a > 0.0);
        CPHI[j] = F.phi[(size_t)mi * L.S_max * (F.phi_row_stride ? L.S_max : 1) + (size_t)i * F.phi_row_stride + j];
        if (rad && a > 0.0) amax = fmax(amax, a);
    }
    amax = wave_max(amax);
    if (lane == 0) amax_w[w] = amax;
    __syncthreads();
    amax = 0.0;
    for (int q = 0; q < nch; q++) amax = fmax(amax, amax_w[q]);
    const double a_ref = amax > 0.0 ? amax : 1.0;
    for (int j = tid; j < S; j += blockDim.x) {
        const size_t c = cbase + (size_t)i * S + j;
        const double a = F.cellArea[c];
        const bool rad = F.radiates ? (F.radiates[c] == 1) : (a > 0.0);
        CAREA[j] = rad ? (float)(a / a_ref) : __builtin_nanf("");
    }
    const double inv_kT = 1.0 / (K_B_OVER_KEV * pow(10.0, F.logT[ro]));
    bool any = false;
    int chain_n = 0;
    auto flush = [&]() {
        int ln = lane, np = N_P;
        asm volatile("" : "+v"(ln));
        asm volatile("" : "+s"(np));
        if (e0 + 64 <= N_E && p0 + 64 <= np) {
            const unsigned boff = (unsigned)((4 * (ln >> 4) * np + (ln & 15)) * 8);
            double* tile = flux + (size_t)e0 * np + p0;
#pragma unroll
            for (int c = 0; c < 4; c++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const double sc = a_ref * __shfl(esc, 16 * c + 4 * (ln >> 4) + r, 64);
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        double* row = tile + (size_t)((16 * c + r) * np + 16 * t);
                        const float v = acc[c][t][r];
                        if (v != 0.f) unsafeAtomicAdd((double*)((char*)row + boff), (double)v * sc);
                    }
                }
        } else {
#pragma unroll
            for (int c = 0; c < 4; c++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int src = 16 * c + 4 * (ln >> 4) + r, ee = e0 + src;
                    const double sc = a_ref * __shfl(esc, src, 64);
                    if (ee >= N_E) continue;
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        const int pq = p0 + 16 * t + (ln & 15);
                        const float v = acc[c][t][r];
                        if (pq < np && v != 0.f) unsafeAtomicAdd(flux + (size_t)ee * np + pq, (double)v * sc);
                    }
                }
        }
#pragma unroll
        for (int c = 0; c < 4; c++)
#pragma unroll
            for (int t = 0; t < 4; t++) acc[c][t] = v4f{0.f, 0.f, 0.f, 0.f};
        any = false;
        chain_n = 0;
    };
    for (int I = 0; I < L.IO_max; I++) {
        const size_t vix = ro * L.IO_max + I;
        if (!F.valid[vix]) break;                            // the whole block
        __syncthreads();   // the previous image's reads of the geometry are done
        const double* g = F.geo + vix * 5 * N_L;
        const bool copy_last = g[5 * N_L - 1] < -1.0;
        for (int q = tid; q < N_L; q += blockDim.x) {
            const int mm = (q == N_L - 1 && copy_last) ? 0 : q;
            const double geom = g[3 * N_L + mm];
            const bool vis = geom != 0.0 && g[4 * N_L + mm] >= -1.0;
            KN[q] = make_double2(vis ? geom : 0.0, 1.0 / g[N_L + mm]);
            PH[q] = g[q];
            if (q < NI) {
                const double wd = g[q + 1] - g[q];
                KH[q] = make_float2((float)wd, (float)(1.0 / wd));
            }
        }
        __syncthreads();
        for (int q = tid; q < NI; q += blockDim.x) {
            bool any_vis = false;
            for (int k = q - 2; k <= q + 2; k++) {
                const int sk = k < 0 ? k + NI : (k >= NI ? k - NI : k);
                any_vis = any_vis || KN[sk].x != 0.0 || KN[sk + 1].x != 0.0;
            }
            LIVE[q] = any_vis ? 1 : 0;
        }
        __syncthreads();
        const double lo = PH[0], hi = PH[N_L - 1];
        // phase lane: rotation point and wrap offsets, as k_wave_fold
        int j0 = 0;
        if (p_ok && S > 0) {
            const double x0 = wrap_phase(ph + CPHI[0], lo, hi);
            int a = 1, b = S;
            while (a < b) {
                const int mid = (a + b) >> 1;
                if (wrap_phase(ph + CPHI[mid], lo, hi) < x0) b = mid;
                else a = mid + 1;
            }
            j0 = a == S ? 0 : a;
        }
        auto cell = [&](int k) { int j = j0 + k; return j >= S ? j - S : j; };
        auto turns = [&](int j) {
            const double s0 = ph + CPHI[j], x = wrap_phase(s0, lo, hi);
            return x == s0 ? 0 : (x == s0 - TWO_PI ? -1 : (x == s0 + TWO_PI ? 1 : 2));
        };
        int tA = 0, tB = 0;
        if (p_ok && S > 0) { tA = turns(j0); tB = j0 > 0 ? turns(0) : 0; }
        const double offA = tA == -1 ? -TWO_PI : (tA == 1 ? TWO_PI : 0.0);
        const double offB = tB == -1 ? -TWO_PI : (tB == 1 ? TWO_PI : 0.0);
        const bool slow = tA == 2 || tB == 2;
        auto xof = [&](int j) -> double {
            const double s0 = ph + CPHI[j];
            double x = s0 + (j >= j0 ? offA : offB);
            if (slow) x = wrap_phase(s0, lo, hi);
            return x;
        };
        // the rotation must be one sorted turn within [lo, hi] (checked once, by wavefront 0)
        if (w == 0) {
            bool unsorted = false;
            if (p_ok) {
                double prev = lo;
                for (int k = 0; k < S; k++) {
                    const double x = xof(cell(k));
                    unsorted = unsorted || x < prev || x > hi;
                    prev = x;
                }
            }
            if (__ballot(unsorted) && lane == 0) members[mi].status = 2;
        }
        const int nlanes = p_ok ? S : 0;   // cells a phase lane walks

        auto knot = [&](double2 kn) -> float {
            if (kn.x == 0.0 || !e_ok) return 0.f;
            const double Ep = E * kn.y;
            const float Ef = (float)Ep;
            const float Ep3 = Ef * Ef * Ef;
            float I_E = Ep3 / (expf((float)(Ep * inv_kT)) - 1.0f);
            if (correct) I_E -= Ep3 / (expf((float)(Ep * inv_kT_else)) - 1.0f);
            return I_E * (float)kn.x;
        };
        const float yl2 = knot(KN[NI - 2]), yl1 = knot(KN[NI - 1]), yl0 = knot(KN[NI]);
        float ya = knot(KN[0]), yb = knot(KN[1]), yk = knot(KN[2]);
        float sm2 = (yl1 - yl2) * KH[NI - 2].y, sm1 = (yl0 - yl1) * KH[NI - 1].y;
        float s0 = (yb - ya) * KH[0].y, s1 = (yk - yb) * KH[1].y;
        const float s0_first = s0, s1_first = s1;
        float w_c = 0.f, al_c = 0.f;
        bool carry = false;
        double2 kn_next = KN[min(3, NI)];

        for (int c0 = 0; c0 < NI; c0 += WF_SLAB) {
            const int nsl = min(WF_SLAB, NI - c0);
            // -- the walk: this wavefront's share of the slab's intervals
            const int iv_a = c0 + (w * nsl) / nch, iv_b = c0 + ((w + 1) * nsl) / nch;
            if (iv_a < iv_b) {
                // first cell (rotation order) at or past the share's first leaf
                int cur = 0;
                if (iv_a > 0) {
                    int a = 0, b = nlanes;
                    const double lb0 = PH[iv_a];
                    while (a < b) {
                        const int mid = (a + b) >> 1;
                        if (xof(cell(mid)) < lb0) a = mid + 1;
                        else b = mid;
                    }
                    cur = a;
                }
                for (int iv = iv_a; iv < iv_b; iv++) {
                    const double lb = PH[iv], ub = iv < NI - 1 ? PH[iv + 1] : __builtin_inf();
                    const bool piece_live = LIVE[iv] != 0;
                    CF[(iv - c0) * 64 + lane] = cur;
                    float M0 = 0.f, M1 = 0.f, M2 = 0.f, M3 = 0.f;
                    int cnt = 0;
                    while (cur < nlanes) {
                        const int j = cell(cur);
                        const double x = xof(j);
                        if (!(x < ub)) break;
                        const float an = CAREA[j];
                        if (piece_live && !isnan(an)) {
                            const float tq = (float)(x - lb);
                            float at = an;
                            M0 += at; at *= tq; M1 += at; at *= tq; M2 += at; at *= tq; M3 += at;
                            cnt++;
                        }
                        cur++;
                    }
                    float* mo = MOM + (iv - c0) * 256;
                    mo[lane] = M0; mo[64 + lane] = M1; mo[128 + lane] = M2; mo[192 + lane] = M3;
                    const unsigned long long nz = __ballot(cnt > 0);
                    if (lane == 0) MASK[iv - c0] = nz;
                }
                if (iv_b == c0 + nsl) CF[nsl * 64 + lane] = cur;
            }
            __syncthreads();
            // -- the slab's intervals for this wavefront's energies
            for (int iv = c0; iv < c0 + nsl; iv++) {
                const int k2 = iv + 2;
                const double2 kn_now = kn_next;
                if (k2 + 2 <= NI) kn_next = KN[k2 + 2];
                float s2, ynew = 0.f;
                if (k2 < NI) {
                    ynew = knot(kn_now);
                    s2 = (ynew - yk) * KH[k2].y;
                } else s2 = k2 == NI ? s0_first : s1_first;
                const bool piece_live = LIVE[iv] != 0;
                const unsigned long long nzB = MASK[iv - c0];
                const float wn = fabsf(s2 - s1) + fabsf(s0 - sm1);
                if (piece_live && nzB) {
                    if (!carry) {
                        w_c = fabsf(s1 - s0) + fabsf(sm1 - sm2);
                        al_c = w_c != 0.f ? fabsf(sm1 - sm2) / w_c : 0.f;
                    }
                    const float h = KH[iv].x, ih = KH[iv].y;
                    const float ca = ya;
                    float cb, cc, cd;
                    const float aln = wn != 0.f ? fabsf(s0 - sm1) / wn : 0.f;
                    if (w_c == 0.f) { cb = s0; cc = 0.f; cd = 0.f; }
                    else {
                        const float tn = wn == 0.f ? s0 : (1.f - aln) * s0 + aln * s1;
                        cb = (1.f - al_c) * sm1 + al_c * s0;
                        cc = (3.f * s0 - 2.f * cb - tn) * ih;
                        cd = (cb + tn - 2.f * s0) * (ih * ih);
                    }
                    w_c = wn;
                    al_c = aln;
                    carry = true;
                    const int cls = correct ? 1 : piece_class(ca, cb, cc, cd, h);
                    float P[4];
                    const bool on = cls == 1 && e_ok;
                    P[0] = on ? ca : 0.f; P[1] = on ? cb : 0.f; P[2] = on ? cc : 0.f; P[3] = on ? cd : 0.f;
                    const unsigned long long nzA =
                        __ballot(on && (ca != 0.f || cb != 0.f || cc != 0.f || cd != 0.f));
                    const unsigned long long mixA = __ballot(cls == 2 && e_ok);
                    if (nzA) {
                        row_transpose(P);
                        // B operands of the four phase tiles, stored in their order
                        const float* mo = MOM + (iv - c0) * 256 + (lane >> 4) * 64 + (lane & 15);
                        float Mt[4];
#pragma unroll
                        for (int t = 0; t < 4; t++) Mt[t] = mo[16 * t];
#pragma unroll
                        for (int c = 0; c < 4; c++)
#pragma unroll
                            for (int t = 0; t < 4; t++)
                                acc[c][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(P[c], Mt[t], acc[c][t], 0, 0, 0);
                        any = true;
                        if (++chain_n == F.chain) flush();
                    }
                    if (mixA) {
                        const int c_first = CF[(iv - c0) * 64 + lane], c_last = CF[(iv - c0 + 1) * 64 + lane];
                        const double lb = PH[iv];
                        unsigned long long mb = mixA;
                        while (mb) {
                            const int el = __ffsll((long long)mb) - 1;
                            mb &= mb - 1ull;
                            const float a = readlane_f(ca, el), b = readlane_f(cb, el), c = readlane_f(cc, el),
                                        d = readlane_f(cd, el);
                            const double sc = a_ref * readlane_d(esc, el);
                            float sum = 0.f;
                            for (int k = c_first; k < c_last; k++) {
                                const int j = cell(k);
                                const float ar = CAREA[j];
                                if (isnan(ar)) continue;
                                const float tq = (float)(xof(j) - lb);
                                const float v = a + tq * (b + tq * (c + tq * d));
                                if (v > 0.f) sum += ar * v;
                            }
                            if (sum != 0.f && p_ok) unsafeAtomicAdd(flux + (size_t)(e0 + el) * N_P + p, (double)sum * sc);
                        }
                    }
                } else carry = false;
                ya = yb; yb = yk; yk = ynew;
                sm2 = sm1; sm1 = s0; s0 = s1; s1 = s2;
            }
            __syncthreads();   // MOM, MASK and CF are rewritten by the next slab
        }
    }
    if (any) flush();
}

}  // namespace xh
