// Moment form of the cell x phase x energy fold, on the matrix cores.
//
// Reference behaviour: the energy x cell x phase spline loop of
// xpsi/cellmesh/integrator_for_azimuthal_invariance.pyx:565-596 and the
// blackbody intensity of xpsi/surface_radiation_field/hot_BB.pyx:54-98.
//
// For one parallel and image order the reference evaluates, for every energy
// k, the periodic Akima interpolant s_k of the profile at x = phase_p + phi_j
// for every radiating cell j and model phase p, keeps values > 0 and adds
// cellArea_j * s_k(x).  The interval i(j, p) holding x and the offset
// t = x - PHASE[i] do not depend on the energy.  Where s_k is positive over
// the whole interval the clipping is inactive and
//
//     sum_j A_j s_k(x_jp) = sum_i sum_m coef_k[i][m] * MOM[i][m][p],
//     MOM[i][m][p]       = sum over cells j with i(j, p) = i of A_j t^m,
//
// the product of an [energy x (interval, power)] matrix with an
// [(interval, power) x phase] matrix: one 16x16x4 MFMA per interval and
// 16 x 16 tile.  Intervals where s_k changes sign are evaluated point by point
// with the clipping, exactly as the reference does; intervals where s_k <= 0
// throughout contribute nothing.
//
// One block of four waves owns one 16-phase tile of one member and up to 256
// energies, and walks the parallels of its ring block.  Per parallel and
// image: the tile's moments from the (cell, phase) queries by LDS atomics
// (energy independent, once per block); then per 16-energy chunk the profile
// (intensity x geometry), secant slopes, Akima coefficients with a sign class
// per (interval, energy), direct evaluation of sign-changing pieces, and the
// MFMA sweep, intervals dealt round-robin to the waves.  Each parallel's
// tile is added into FP64 accumulators, so the float form sums parallels in
// FP64.
//
// R = double: FP64 throughout (v_mfma_f64_16x16x4f64).
// R = float: the precision form.  Blackbody intensity, spline coefficients,
// moments and products in FP32 (v_mfma_f32_16x16x4f32); the intensity's
// arguments (E' = E/Z and E'/kT) and the phase offsets t are formed in FP64
// and rounded once, the per-parallel partial sums are accumulated in FP64 and
// the blackbody norm is applied in FP64 at the end.
#include "xh_common.hpp"
#include "xh_types.hpp"
#include "xh_kernels.hpp"
#include "xh_atm.hpp"

namespace xh {

template <typename R> struct Mfma;

// lane l: A[l % 16][l / 16], B[l / 16][l % 16]; result element r of lane l is
// D[4 r + l / 16][l % 16]
template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 step(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return 4 * r + (l >> 4); }
};

// lane l: A[l % 16][l / 16], B[l / 16][l % 16]; result element r of lane l is
// D[4 (l / 16) + r][l % 16]
template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 step(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};

template <typename R> struct Eps;
template <> struct Eps<double> { static constexpr double v = DBL_EPSILON; };
template <> struct Eps<float> { static constexpr float v = FLT_EPSILON; };

// Sign class of f(t) = a + t (b + t (c + t d)) on [0, h], with the
// evaluation-error margin tol:
// 1 = f >= -tol on the piece: the clipping can only drop values within the
//     rounding margin of zero (pieces that end on a zero knot with zero
//     slope, next to hidden leaves, fall here);
// 0 = f <= tol on the piece, or identically zero;
// 2 = otherwise, evaluated point by point with the clipping.
// The cubic Bernstein coefficients bound f (the curve lies in their convex
// hull) and settle most pieces; the rest are decided by f at the ends and at
// the critical points in (0, h).  The classes 0 and 1 differ from per-point
// clipping by at most tol x area per query, at queries whose value lies
// within tol of zero.
template <typename R>
__device__ __forceinline__ int cubic_class(R a, R b, R c, R d, R h)
{
    if (a == (R)0 && b == (R)0 && c == (R)0 && d == (R)0) return 0;
    const R bh = b * h, ch = c * h * h, dh = d * h * h * h;
    const R scale = fabs(a) + fabs(bh) + fabs(ch) + fabs(dh);
    const R tol = (R)64 * Eps<R>::v * scale;
    const R b0 = a, b1 = a + bh / (R)3, b2 = a + (R)2 * bh / (R)3 + ch / (R)3, b3 = a + bh + ch + dh;
    R lo = fmin(fmin(b0, b1), fmin(b2, b3)), hi = fmax(fmax(b0, b1), fmax(b2, b3));
    if (lo >= -tol) return 1;
    if (hi <= tol) return 0;
    // extremes: the ends and the roots of f'(t) = b + 2 c t + 3 d t^2 in (0, h)
    auto f = [&](R t) { return a + t * (b + t * (c + t * d)); };
    lo = fmin(a, b3);
    hi = fmax(a, b3);
    R r1 = (R)-1, r2 = (R)-1;
    if (d != (R)0) {
        const R disc = c * c - (R)3 * b * d;
        if (disc >= (R)0) {
            const R q = -(c + copysign(sqrt(disc), c));
            if (q != (R)0) { r1 = q / ((R)3 * d); r2 = b / q; }
            else r1 = (R)0;
        }
    } else if (c != (R)0) r1 = -b / ((R)2 * c);
    if (r1 > (R)0 && r1 < h) { const R v = f(r1); lo = fmin(lo, v); hi = fmax(hi, v); }
    if (r2 > (R)0 && r2 < h) { const R v = f(r2); lo = fmin(lo, v); hi = fmax(hi, v); }
    if (lo >= -tol) return 1;
    if (hi <= tol) return 0;
    return 2;
}

// The (cell, phase) queries of one thread: cells ja .. jb-1 of a parallel at
// model phase ph, in azimuth order.  Each radiating cell's point
// x = ph + phi_j is wrapped into [lo, hi] as the reference wraps it, its
// interval cur (PH[cur] <= x < PH[cur + 1], the last interval closed at hi)
// found by stepping on from the previous cell's, and
// visit(cur, x - PH[cur], area, entered) called, entered when cur differs
// from the previous radiating cell's.  The interval bounds stay in registers
// and the next cell is read one iteration ahead.
template <typename R, typename V>
__device__ __forceinline__ void walk_queries(const double* __restrict__ PH, int N_L,
                                             const double* __restrict__ cphi,
                                             const R* __restrict__ carea, int ja, int jb, double ph,
                                             V&& visit)
{
    const int NI = N_L - 1;
    const double lo = PH[0], hi = PH[N_L - 1];
    int cur = -1;
    double blo = 1.0, bhi = 0.0;   // empty: the first cell searches
    double cpn = ja < jb ? cphi[ja] : 0.0;
    R can = ja < jb ? carea[ja] : (R)0;
#pragma unroll 1
    for (int j = ja; j < jb; j++) {
        const double cpj = cpn;
        const R caj = can;
        if (j + 1 < jb) { cpn = cphi[j + 1]; can = carea[j + 1]; }
        if (isnan(caj)) continue;
        const double x = wrap_phase(ph + cpj, lo, hi);
        bool entered = false;
        if (!(x >= blo && x < bhi)) {
            int id = cur;
            if (id < 0 || x < blo) id = interval_of(PH, N_L, x);
            else while (id < NI - 1 && PH[id + 1] <= x) id++;
            cur = id;
            blo = PH[id];
            bhi = id < NI - 1 ? PH[id + 1] : __builtin_inf();
            entered = true;
        }
        visit(cur, x - blo, caj, entered);
    }
}

// Block-wide stream compaction of NS flag sets at once (one value per thread,
// up to 256 per set) in thread order; cnt holds NS NW scratch ints.  Every
// thread of the block calls it (it holds two barriers).  Writes the counts to
// n[0 .. NS-1].
template <int NW, int NS>
__device__ __forceinline__ void compact_u16(const bool (&f)[NS], int value, unsigned short* const (&o)[NS], int* cnt,
                                            int (&n)[NS])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long b[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
        b[k] = __ballot(f[k]);
        if (lane == 0) cnt[k * NW + wave] = __popcll(b[k]);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NS; k++) {
        int base = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            const int c = cnt[k * NW + w];
            base += w < wave ? c : 0;
            tot += c;
        }
        if (f[k]) o[k][base + __popcll(b[k] & below)] = (unsigned short)value;
        n[k] = tot;
    }
    __syncthreads();
}

// One block of eight waves owns T 16-phase tiles of one member, the cells
// of its ring block and moment_groups(T) groups of G 16-energy chunks, one
// group resident at a time.  Per parallel and image: the T tiles' moments from the (cell, phase)
// queries (energy independent, once), which also mark the intervals each tile
// reaches; of those, the intervals whose Akima stencil holds a visible leaf
// are listed (the others fold a zero piece).  Then per group of G chunks, for
// the listed intervals and the knots and secants their stencils use: the
// profile, Akima coefficients and sign classes, a direct pass over the
// queries for the sign-changing pieces (only when the group has any), and
// the MFMA sweeps.  With T G >= 8 pairs wave w owns the (tile, chunk) pairs
// w, w + 8, ... over all listed intervals; with fewer, the 8 / (T G) waves of
// a pair share its intervals.  Each wave adds each parallel's sweeps into
// FP64 accumulators in registers; a pair's waves are summed at the end.
// OCC = waves per SIMD the register budget is sized for: 4 when two blocks
// fit in LDS (the occupancy outweighs the spills the cap costs), else 2.
template <typename R, int ATM, int T, int G, int OCC>
__global__ void __launch_bounds__(512, OCC) k_moment_fold(const Member* __restrict__ members, Layout L, FoldArgs F)
{
    constexpr int NT = 512, NW = 8, P = T * G;
    constexpr int PPW = P >= NW ? P / NW : 1, WPP = P >= NW ? 1 : NW / P;   // pairs per wave, waves per pair
    constexpr int PT = 16 * T, EG = 16 * G, NGB = moment_groups(T);
    static_assert(P % NW == 0 || NW % P == 0, "pairs over waves");
    extern __shared__ __align__(16) double smd[];
    const int mi = blockIdx.z, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const Member m = members[mi];
    if (!m.active || m.status) return;
    const int S = m.S, N_L = L.N_L, N_P = L.N_P, N_E = L.N_E, NI = N_L - 1;
    const int epb = EG * NGB;               // energies per block
    const int n_eb = (N_E + epb - 1) / epb;
    const int eb0 = (blockIdx.x % n_eb) * epb, p0 = (blockIdx.x / n_eb) * PT;
    const int ng = min(NGB, (N_E - eb0 + EG - 1) / EG);
    const int r0 = blockIdx.y * F.ring_block;
    if (r0 >= S) return;
    const int r1 = min(S, r0 + F.ring_block);
    const int pair0 = P >= NW ? wave : wave % P, sub = P >= NW ? 0 : wave / P;

    double* PH = smd;                       // [N_L] leaf phases
    double* Zi = PH + N_L;                  // [N_L] 1 / redshift factor
    double* As = Zi + N_L;                  // [N_L] mu * eta (table atmosphere only)
    double* Gs = As + (ATM == 2 ? N_L : 0); // [N_L] geometric factor
    double* cphi = Gs + N_L;                // [S_max]
    double* ST = cphi + L.S_max;            // [8] previous table query; area slots
    R* carea = (R*)(ST + 8);                // [S_max], NaN = not radiating
    unsigned* MIX = (unsigned*)(carea + L.S_max);   // [G][N_L] energy bits of sign-changing pieces
    unsigned* HITW = MIX + G * N_L;         // [(N_L + 7) / 8] 4 bits per interval: tile t reaches it
    int* CNT = (int*)(HITW + ((N_L + 7) >> 3));   // [64] compaction scratch (56), flag (60)
    short* PV = (short*)(CNT + 64);         // [N_L] previous queried leaf (-2: not queried)
    unsigned short* IL = (unsigned short*)(PV + N_L);   // [N_L] intervals folded
    unsigned short* SLL = IL + N_L;                       // [N_L] secants formed
    unsigned short* KL = SLL + N_L;                       // [N_L] knots evaluated
    unsigned short* ILT = KL + N_L;                       // [T][N_L] folded intervals tile t reaches
    // (offsets from smd keep the LDS address space visible to the compiler)
    R* CO = (R*)((char*)smd + ((((char*)(ILT + T * N_L) - (char*)smd) + 15) & ~(ptrdiff_t)15));   // [G][N_L][4][16]
    R* MOM = CO + G * N_L * 64;             // [T][NI][4][16] moment m at phase p of tile t
    R* SL = MOM + T * NI * 64;              // [G][NI][16] secant slopes, then:
    R* DIR = SL;                            // [G][16][PT] direct sums of sign-changing pieces
    int* MIXED = CNT + 60;                  // the group has sign-changing pieces (set, then read after a barrier)

    // FP32 form: cell areas enter the moments relative to the ring's largest
    // one.  The mesh is equal-area, so interior cells become exactly 1 and
    // their common area multiplies each parallel's sums in FP64 (one FP32
    // rounding of it would scale every bin alike).  Two slots alternate with
    // the parallel's parity.
    unsigned long long* ABITS = (unsigned long long*)(ST + 6);
    if (tid < 2) ABITS[tid] = 0ull;
    double tot[PPW][NGB][4];                // the wave's share of its pairs, per chunk group
#pragma unroll
    for (int u = 0; u < PPW; u++)
#pragma unroll
        for (int g = 0; g < NGB; g++)
#pragma unroll
            for (int r = 0; r < 4; r++) tot[u][g][r] = 0.0;
    const size_t cbase = (size_t)mi * L.S_max * L.S_max;
    // query role: phase pp of the block, cells ja .. jb-1 of the parallel
    const int pp = tid % PT, part = tid / PT, nparts = NT / PT;
    const int ja = (part * S) / nparts, jb = ((part + 1) * S) / nparts;
    const int pg = p0 + pp;
    const double ph = pg < N_P ? F.phases[pg] : 0.0;
    // profile role: energy tid % EG of each group
    double Eth[NGB];
#pragma unroll
    for (int g = 0; g < NGB; g++) Eth[g] = F.energies[min(eb0 + g * EG + tid % EG, N_E - 1)];

    for (int i = r0; i < r1; i++) {
        const size_t ro = (size_t)mi * L.S_max + i;
        __syncthreads();   // the previous parallel is done with the cell arrays
        if (tid == 0) ABITS[(i + 1) & 1] = 0ull;
        for (int j = tid; j < S; j += NT) {
            size_t c = cbase + (size_t)i * S + j;
            double a = F.cellArea[c];
            bool rad = F.radiates ? (F.radiates[c] == 1) : (a > 0.0);
            cphi[j] = F.phi[(size_t)mi * L.S_max * (F.phi_row_stride ? L.S_max : 1) +
                            (size_t)i * F.phi_row_stride + j];
            if (sizeof(R) == 4 && rad && a > 0.0) atomicMax(&ABITS[i & 1], (unsigned long long)__double_as_longlong(a));
        }
        __syncthreads();   // the ring's largest cell area is complete
        const double a_ref = (sizeof(R) == 4 && ABITS[i & 1]) ? __longlong_as_double((long long)ABITS[i & 1]) : 1.0;
        for (int j = tid; j < S; j += NT) {
            size_t c = cbase + (size_t)i * S + j;
            double a = F.cellArea[c];
            bool rad = F.radiates ? (F.radiates[c] == 1) : (a > 0.0);
            carea[j] = rad ? (R)(a / a_ref) : (R)__builtin_nan("");
        }
        const double kT = K_B_OVER_KEV * pow(10.0, F.logT[ro]);
        const double kT_else = K_B_OVER_KEV * pow(10.0, m.else_logT);
        const double p10 = ATM == 2 ? pow(10.0, 3.0 * F.logT[ro]) : 0.0;   // the table's 10^(3 log T)
        for (int I = 0; I < L.IO_max; I++) {
            const size_t vix = ro * L.IO_max + I;
            if (!F.valid[vix]) break;
            __syncthreads();   // the previous image is done with the geometry, lists, CO and MOM
            const double* g = F.geo + vix * 5 * N_L;
            for (int q = tid; q < N_L; q += NT) {
                PH[q] = g[q]; Zi[q] = 1.0 / g[N_L + q]; Gs[q] = g[3 * N_L + q];
                if (ATM == 2) As[q] = g[2 * N_L + q];
                PV[q] = (short)g[4 * N_L + q];
            }
            for (int q = tid; q < ((N_L + 7) >> 3); q += NT) HITW[q] = 0u;
            for (int q = tid; q < G * N_L; q += NT) MIX[q] = 0u;
            for (int q = tid; q < T * NI * 64; q += NT) MOM[q] = (R)0;
            if (ATM == 2 && tid == 0) prev_query_state(F, L, mi, i, I, ST);
            __syncthreads();
            // moments of the tiles' queries (energy independent); reached intervals
            // first and last interval this thread's cells reach (in walk order,
            // wrapping at most once): the direct pass below visits only
            // threads whose range holds a sign-changing piece
            int qfirst = -1, qlast = -1;
            if (pg < N_P) {
                const int t = pp >> 4, col = pp & 15;
                int mcur = -1;
                R m0 = (R)0, m1 = (R)0, m2 = (R)0, m3 = (R)0;
                auto flush = [&]() {
                    R* mo = MOM + (t * NI + mcur) * 64 + col;
                    atomicAdd(mo, m0);
                    atomicAdd(mo + 16, m1);
                    atomicAdd(mo + 32, m2);
                    atomicAdd(mo + 48, m3);
                    atomicOr(&HITW[mcur >> 3], 1u << (((mcur & 7) << 2) + t));
                };
                walk_queries(PH, N_L, cphi, carea, ja, jb, ph, [&](int cur, double tq64, R area, bool entered) {
                    if (entered) {
                        if (mcur >= 0) flush();
                        mcur = cur;
                        if (qfirst < 0) qfirst = cur;
                        qlast = cur;
                        m0 = (R)0; m1 = (R)0; m2 = (R)0; m3 = (R)0;
                    }
                    const R tq = (R)tq64;
                    R at = area;
                    m0 += at; at *= tq; m1 += at; at *= tq; m2 += at; at *= tq; m3 += at;
                });
                if (mcur >= 0) flush();
            }
            __syncthreads();
            // a piece folds a nonzero profile only when a knot of its stencil
            // (secants i-2 .. i+2, periodic) shows the leaf
            auto visible = [&](int k) {
                if (k == N_L - 1 && PV[k] < -1) k = 0;   // the periodic copy of the first leaf
                return Gs[k] != 0.0 && PV[k] >= -1;
            };
            auto hits = [&](int q) { return (HITW[q >> 3] >> ((q & 7) << 2)) & 15u; };
            bool keep = false;
            if (tid < NI && hits(tid)) {
                for (int d = -2; d <= 3; d++) {
                    int k = tid + d;
                    k = k < 0 ? k + NI : (k > NI ? k - NI : k);
                    keep = keep || visible(k);
                }
            }
            __syncthreads();
            if (tid < NI && !keep) atomicAnd(&HITW[tid >> 3], ~(15u << ((tid & 7) << 2)));
            __syncthreads();
            // lists: folded intervals; secants (i-2 .. i+2) and knots
            // (i-2 .. i+3) of their stencils, the two end knots always
            bool fsec = false, fknot = tid == 0 || tid == N_L - 1;
            if (tid < NI) {
                for (int d = -3; d <= 2; d++) {
                    int q = tid + d;
                    q = q < 0 ? q + NI : (q >= NI ? q - NI : q);
                    if (hits(q)) {
                        fknot = true;
                        if (d >= -2) fsec = true;
                    }
                }
            }
            // lists: folded intervals, secants, knots, and per tile the folded
            // intervals it reaches (the others hold zero moments for it)
            bool flags[3 + T];
            unsigned short* outs[3 + T];
            int ncount[3 + T];
            flags[0] = tid < NI && hits(tid) != 0u; outs[0] = IL;
            flags[1] = fsec; outs[1] = SLL;
            flags[2] = tid < N_L && fknot; outs[2] = KL;
#pragma unroll
            for (int t = 0; t < T; t++) {
                flags[3 + t] = tid < NI && ((hits(tid) >> t) & 1u);
                outs[3 + t] = ILT + t * N_L;
            }
            compact_u16<NW, 3 + T>(flags, tid, outs, CNT, ncount);
            const int n_i = ncount[0], n_s = ncount[1], n_k = ncount[2];
            if (n_i == 0) continue;
            const bool copy_last = PV[N_L - 1] < -1;   // the last leaf repeats the first
            for (int gi = 0; gi < ng; gi++) {
                const int e0 = eb0 + gi * EG;
                __syncthreads();   // the previous group's sweeps are done
                if (tid == 0) *MIXED = 0;
                for (int q = tid; q < n_i * G; q += NT) MIX[(q % G) * N_L + IL[q / G]] = 0u;
                // profile: specific intensity x geometry, zero where the leaf is hidden
                const double Ee = Eth[gi];
                // the table atmosphere for a query with no predecessor on its own leaf
                // (a leaf's first energy): the reference's base-node history in full
                auto table_direct = [&](int mm, int e, double Ep) -> double {
                    double vec[4], prev[4];
                    int has_prev = 1;
                    vec[0] = F.logT[ro]; vec[1] = F.logg[ro]; vec[2] = As[mm];
                    vec[3] = log10(Ep / kT);
                    if (e > 0) {
                        prev[0] = vec[0]; prev[1] = vec[1]; prev[2] = vec[2];
                        prev[3] = log10((F.energies[e - 1] * Zi[mm]) / kT);
                    } else if (PV[mm] >= 0) {
                        int kd = PV[mm];
                        prev[0] = vec[0]; prev[1] = vec[1]; prev[2] = As[kd];
                        prev[3] = log10((F.energies[N_E - 1] * Zi[kd]) / kT);
                    } else if (ST[0] > 0.5) {
                        prev[0] = ST[1]; prev[1] = ST[2]; prev[2] = ST[3]; prev[3] = ST[4];
                    } else has_prev = 0;
                    return num4d_I(F, vec, prev, has_prev);
                };
                if (ATM == 2) {
                    // runs of eight consecutive energies of one leaf: a leaf's later
                    // energies from the parallel's contracted table (their predecessor
                    // query is the same leaf's previous energy, so log T, log g and mu
                    // keep their base nodes and never clamp); a leaf's first energy
                    // queries the table with the full history
                    constexpr int RUN = 8;
                    constexpr int NRUN = EG / RUN;
                    const double* slab = (F.slab && F.atm_S) ? F.slab + ro * (size_t)F.atm_N[2] * F.atm_N[3] : nullptr;
                    for (int q = tid; q < n_k * NRUN; q += NT) {
                        const int ms = KL[q / NRUN], mm = (ms == N_L - 1 && copy_last) ? 0 : ms;
                        const int k0 = (q % NRUN) * RUN;
                        const double geom = Gs[mm], zi = Zi[mm];
                        const bool vis = geom != 0.0 && PV[mm] >= -1;
                        double Iv[RUN];
                        int ks = k0;   // first energy of the run served by the slab
                        if (vis && slab) {
                            if (e0 + ks == 0) ks++;
                            const int n = min(k0 + RUN, N_E - e0) - ks;
                            if (n > 0) {
                                double qv[RUN];
#pragma unroll
                                for (int k = 0; k < RUN; k++)
                                    qv[k] = k < n ? log10((F.energies[e0 + ks + k] * zi) / kT) : 0.0;
                                const double qp = log10((F.energies[e0 + ks - 1] * zi) / kT);
                                n4_slab_run<RUN>(F, slab, As[mm], qv, qp, n, p10, Iv + (ks - k0));
                            }
                        }
#pragma unroll
                        for (int k = 0; k < RUN; k++) {
                            const int ke = k0 + k, e = e0 + ke;
                            R v = (R)0;
                            if (vis && e < N_E) {
                                const double Ep = F.energies[e] * zi;
                                const double I = (slab && ke >= ks) ? Iv[k] : table_direct(mm, e, Ep);
                                double corr = 0.0;
                                if (m.correct) corr = Ep * Ep * Ep / (exp(Ep / kT_else) - 1.0) * BB_NORM;
                                v = (R)((I * NUM4D_NORM - corr) * geom);
                            }
                            CO[((ke >> 4) * N_L + ms) * 64 + (ke & 15)] = v;
                        }
                    }
                } else {
                for (int q = tid; q < n_k * EG; q += NT) {
                    const int ms = KL[q / EG], mm = (ms == N_L - 1 && copy_last) ? 0 : ms;
                    const int ke = q % EG, e = e0 + ke;
                    R v = (R)0;
                    const double geom = Gs[mm];
                    if (geom != 0.0 && e < N_E && PV[mm] >= -1) {
                        const double Ep = Ee * Zi[mm];
                        if (sizeof(R) == 8) {
                            double I_E = Ep * Ep * Ep / (exp(Ep / kT) - 1.0);
                            double corr = 0.0;
                            if (m.correct) corr = Ep * Ep * Ep / (exp(Ep / kT_else) - 1.0) * BB_NORM;
                            v = (R)((I_E * BB_NORM - corr) * geom);
                        } else {
                            // FP32 intensity from arguments rounded once each: a shared
                            // rounding (of E, kT or the norm) would bias every bin alike,
                            // so E' and E'/kT are formed in FP64 and the norm is applied
                            // in FP64 after the sums
                            const float Ef = (float)Ep;
                            const float Ep3 = Ef * Ef * Ef;
                            float I_E = Ep3 / (expf((float)(Ep / kT)) - 1.0f);
                            if (m.correct) I_E -= Ep3 / (expf((float)(Ep / kT_else)) - 1.0f);
                            v = (R)(I_E * (float)geom);
                        }
                    }
                    CO[((ke >> 4) * N_L + ms) * 64 + (ke & 15)] = v;
                }
                }
                __syncthreads();
                // secant slopes
                for (int q = tid; q < n_s * EG; q += NT) {
                    const int mm = SLL[q / EG], ke = q % EG, gg = ke >> 4, k = ke & 15;
                    const R* cg = CO + gg * N_L * 64;
                    SL[(gg * NI + mm) * 16 + k] = (cg[(mm + 1) * 64 + k] - cg[mm * 64 + k]) / (R)(PH[mm + 1] - PH[mm]);
                }
                __syncthreads();
                // periodic Akima coefficients and the sign class of each piece
                for (int q = tid; q < n_i * EG; q += NT) {
                    const int mm = IL[q / EG], ke = q % EG, gg = ke >> 4, k = ke & 15;
                    const R* sg = SL + gg * NI * 16;
                    auto sl = [&](int kk) {
                        kk = kk < 0 ? kk + NI : (kk >= NI ? kk - NI : kk);
                        return sg[kk * 16 + k];
                    };
                    R* co = CO + (gg * N_L + mm) * 64 + k;
                    const R h = (R)(PH[mm + 1] - PH[mm]);
                    R a = co[0], b, cc, d;
                    akima_piece_t<R>(sl(mm - 2), sl(mm - 1), sl(mm), sl(mm + 1), sl(mm + 2), h, b, cc, d);
                    // with the elsewhere correction every value is kept (:593)
                    const int cls = m.correct ? 1 : cubic_class<R>(a, b, cc, d, h);
                    if (cls == 0) { a = (R)0; b = (R)0; cc = (R)0; d = (R)0; }
                    else if (cls == 2) { atomicOr(&MIX[gg * N_L + mm], 1u << k); *MIXED = 1; }
                    co[0] = a;
                    co[16] = b;
                    co[32] = cc;
                    co[48] = d;
                }
                __syncthreads();
                const int mixed = *MIXED;
                if (mixed) {
                    // sign-changing pieces, point by point with the clipping
                    for (int q = tid; q < G * 16 * PT; q += NT) DIR[q] = (R)0;
                    __syncthreads();
                    bool need = false;
                    if (pg < N_P && qfirst >= 0) {
                        int id = qfirst;
                        for (int n = 0; n < NI && !need; n++) {
                            unsigned any = 0u;
#pragma unroll
                            for (int gg = 0; gg < G; gg++) any |= MIX[gg * N_L + id];
                            need = any != 0u;
                            if (id == qlast) break;
                            id = id + 1 == NI ? 0 : id + 1;
                        }
                    }
                    if (need) {
                        unsigned mx[G];
#pragma unroll
                        for (int gg = 0; gg < G; gg++) mx[gg] = 0u;
                        walk_queries(PH, N_L, cphi, carea, ja, jb, ph, [&](int cur, double tq64, R ar, bool entered) {
                            if (entered)
#pragma unroll
                                for (int gg = 0; gg < G; gg++) mx[gg] = MIX[gg * N_L + cur];
                            const R tq = (R)tq64;
#pragma unroll
                            for (int gg = 0; gg < G; gg++) {
                                unsigned mb = mx[gg];
                                while (mb) {
                                    const int k = __ffs(mb) - 1;
                                    mb &= mb - 1u;
                                    const R* co = CO + (gg * N_L + cur) * 64 + k;
                                    const R v = co[0] + tq * (co[16] + tq * (co[32] + tq * co[48]));
                                    if (v > (R)0) atomicAdd(&DIR[(gg * 16 + k) * PT + pp], ar * v);
                                }
                            }
                        });
                    }
                    __syncthreads();
                }
                // the sweeps of the wave's pairs over the folded intervals its
                // tile reaches: the pairs share the tile (pair0 + u * 8 has
                // tile pair0 % T), so one B operand feeds PPW independent
                // accumulator chains; sign-changing pieces are masked out of A
                {
                    // alternate intervals feed two chains per pair (A, B), summed
                    // in FP64: shorter FP32 accumulations and two MFMAs in flight
                    const int wt = pair0 % T;
                    typename Mfma<R>::v4 swA[PPW], swB[PPW];
                    bool on[PPW];
                    const R* cg[PPW];
                    const unsigned* mxg[PPW];
#pragma unroll
                    for (int u = 0; u < PPW; u++) {
                        const int wg = (pair0 + u * NW) / T;
                        on[u] = e0 + wg * 16 < N_E;
                        cg[u] = CO + wg * N_L * 64;
                        mxg[u] = MIX + wg * N_L;
                        swA[u] = typename Mfma<R>::v4{(R)0, (R)0, (R)0, (R)0};
                        swB[u] = swA[u];
                    }
                    const int kb = lane & 15;
                    const R* mt = MOM + wt * NI * 64;
                    auto sweep = [&](int idx, typename Mfma<R>::v4* sw) {
                        const R bm = mt[idx * 64 + lane];
#pragma unroll
                        for (int u = 0; u < PPW; u++) {
                            R a = cg[u][idx * 64 + lane];
                            if ((mxg[u][idx] >> kb) & 1u) a = (R)0;
                            sw[u] = Mfma<R>::step(a, bm, sw[u]);
                        }
                    };
                    // the tile's own list: no branch in the loop
                    int n_t = ncount[3];
#pragma unroll
                    for (int t = 1; t < T; t++)
                        if (wt == t) n_t = ncount[3 + t];
                    const unsigned short* lt = ILT + wt * N_L;
                    if (on[0]) {
                        // software-pipelined: the operands of the next two
                        // intervals are read while this pair's MFMAs run
                        struct Ops { R b; R a[PPW]; };
                        auto fetch = [&](int idx, Ops& o) {
                            o.b = mt[idx * 64 + lane];
#pragma unroll
                            for (int u = 0; u < PPW; u++) {
                                R a = cg[u][idx * 64 + lane];
                                if ((mxg[u][idx] >> kb) & 1u) a = (R)0;
                                o.a[u] = a;
                            }
                        };
                        int q = sub;
                        if (q + WPP < n_t) {
                            Ops c0, c1;
                            fetch(lt[q], c0);
                            fetch(lt[q + WPP], c1);
                            for (; q + 3 * WPP < n_t; q += 2 * WPP) {
                                Ops n0, n1;
                                fetch(lt[q + 2 * WPP], n0);
                                fetch(lt[q + 3 * WPP], n1);
#pragma unroll
                                for (int u = 0; u < PPW; u++) {
                                    swA[u] = Mfma<R>::step(c0.a[u], c0.b, swA[u]);
                                    swB[u] = Mfma<R>::step(c1.a[u], c1.b, swB[u]);
                                }
                                c0 = n0;
                                c1 = n1;
                            }
#pragma unroll
                            for (int u = 0; u < PPW; u++) {
                                swA[u] = Mfma<R>::step(c0.a[u], c0.b, swA[u]);
                                swB[u] = Mfma<R>::step(c1.a[u], c1.b, swB[u]);
                            }
                            q += 2 * WPP;
                        }
                        if (q < n_t) sweep(lt[q], swA);
                    }
#pragma unroll
                    for (int u = 0; u < PPW; u++) {
                        if (!on[u]) continue;
                        const int wg = (pair0 + u * NW) / T;
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            double v = (double)swA[u][r] + (double)swB[u][r];
                            if (mixed && sub == 0)
                                v += (double)DIR[(wg * 16 + Mfma<R>::row(lane, r)) * PT + wt * 16 + (lane & 15)];
                            tot[u][gi][r] += v * a_ref;
                        }
                    }
                }
            }
        }
    }
    // a pair's waves summed; outputs (energy e, phase p)
    double* RED = (double*)CO;              // [8 waves][4][64]; the host sizes LDS to hold it
    for (int gi = 0; gi < ng; gi++) {
        if (WPP > 1) {
            __syncthreads();
#pragma unroll
            for (int r = 0; r < 4; r++) RED[(wave * 4 + r) * 64 + lane] = tot[0][gi][r];
            __syncthreads();
        }
        if (sub != 0) continue;
#pragma unroll
        for (int u = 0; u < PPW; u++) {
            const int pair = pair0 + u * NW, wt = pair % T, wg = pair / T;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                double s = tot[u][gi][r];
                for (int w = wave + P; w < NW; w += P) s += RED[(w * 4 + r) * 64 + lane];
                const int e = eb0 + gi * EG + wg * 16 + Mfma<R>::row(lane, r), pq = p0 + wt * 16 + (lane & 15);
                if (e < N_E && pq < N_P) {
                    if (sizeof(R) == 4 && ATM == 1) s *= BB_NORM;
                    const double E = F.energies[e];
                    double* fo = F.flux + (size_t)mi * N_E * N_P + (size_t)e * N_P + pq;
                    if (gridDim.y == 1) *fo = s / (E * KEV);
                    else atomicAdd(fo, s * (1.0 / (E * KEV)));
                }
            }
        }
    }
}

// The table contracted over (log T, log g) for every contributing parallel:
// slab[member][ring][mu node][energy node], FP64.  Blocks (x) share out one
// parallel's mu x energy plane; consecutive threads read consecutive energy
// nodes of the 16 (log T, log g) rows.
__global__ void __launch_bounds__(256) k_table_slab(const Member* __restrict__ members, Layout L, FoldArgs F,
                                                    double* __restrict__ slab)
{
    const int i = blockIdx.y, mi = blockIdx.z;
    const Member& m = members[mi];
    if (!m.active || m.status || i >= m.S) return;
    const size_t ro = (size_t)mi * L.S_max + i;
    if (!F.valid[ro * L.IO_max]) return;   // the parallel folds nothing
    int b0, b1;
    double w0[4], w1[4];
    n4_ring_weights(F, F.logT[ro], F.logg[ro], b0, b1, w0, w1);
    const size_t s1 = (size_t)F.atm_N[2] * F.atm_N[3], s0 = (size_t)F.atm_N[1] * s1;
    const double* T = F.atm_I + (size_t)b0 * s0 + (size_t)b1 * s1;
    double* out = slab + ro * s1;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < s1; k += (size_t)gridDim.x * blockDim.x) {
        double c = 0.0;
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int b = 0; b < 4; b++) c += w0[a] * w1[b] * T[(size_t)a * s0 + (size_t)b * s1 + k];
        out[k] = c;
    }
}

#define XH_MOMENT_INST(R, T, OCC)                                                                \
    template __global__ void k_moment_fold<R, 1, T, T, OCC>(const Member*, Layout, FoldArgs);   \
    template __global__ void k_moment_fold<R, 2, T, T, OCC>(const Member*, Layout, FoldArgs);
XH_MOMENT_INST(double, 1, 2) XH_MOMENT_INST(double, 2, 2) XH_MOMENT_INST(double, 4, 2)
XH_MOMENT_INST(double, 1, 4) XH_MOMENT_INST(double, 2, 4)
XH_MOMENT_INST(float, 1, 2) XH_MOMENT_INST(float, 2, 2) XH_MOMENT_INST(float, 4, 2)
XH_MOMENT_INST(float, 1, 4) XH_MOMENT_INST(float, 2, 4)

}  // namespace xh
