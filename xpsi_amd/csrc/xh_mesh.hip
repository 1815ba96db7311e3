// Hot-region cell mesh on the GPU.
//
// Reference behaviour: xpsi/cellmesh/mesh_tools.pyx:18-1099 (allocate_cells,
// area/spot/cell integrals), mesh.pyx:23-417 (spot mesh), polar_mesh.pyx:
// 23-430 (polar-cap mesh), effective_gravity_universal.pyx:10-71.
// Every adaptive integral is one wavefront (wave_qag61); the 1000-point
// equal-area table is 1000 independent waves; each ring is one wave whose
// lanes classify the ring's cells and then integrate boundary cells in turn.
#include "xh_common.hpp"
#include "xh_types.hpp"
#include "xh_kernels.hpp"

namespace xh {

__device__ __forceinline__ double radius_norm(double mu, double eps, double zeta, int shape)
{
    return shape == 0 ? 1.0 + eps * (-0.788 + 1.030 * zeta) * mu * mu : 1.0;
}

__device__ __forceinline__ double f_theta(double mu, double rn, double eps, double zeta, int shape)
{
    double d = -2.0 * eps * (-0.788 + 1.030 * zeta) * mu * sqrt(1.0 - mu * mu);
    return shape == 0 ? d / (rn * sqrt(1.0 - 2.0 * zeta / rn)) : 0.0;
}

__device__ double effective_gravity(double mu, double R_eq, double x, double eps, int shape)
{
    double g0 = x * C_SI * C_SI / (R_eq * sqrt(1.0 - 2.0 * x));
    double e2 = eps * eps;
    double ce = -0.791 + 0.776 * x, cp = 1.138 - 1.431 * x;
    double de = (-1.315 + 2.431 * x) * eps * x;
    double dp = (0.653 - 2.864 * x) * eps * x;
    double d60 = (13.47 - 27.13 * x) * eps * x;
    double fe = -1.172 * x * e2, fp = 0.975 * x * e2;
    double g = 1.0;
    g += (ce + de + fe) * eps * (1.0 - mu * mu);
    g += (cp + dp + fp - d60) * eps * mu * mu;
    g += d60 * eps * fabs(mu);
    return shape == 0 ? log10(g * g0) + 2.0 : log10(g0) + 2.0;
}

struct AreaFn {
    double eps, zeta;
    int av, shape;
    __device__ double operator()(double th) const
    {
        if (eq12(th, 0.0)) return 0.0;
        double mu = cos(th);
        double r = radius_norm(mu, eps, zeta, shape);
        double f = f_theta(mu, r, eps, zeta, shape);
        double v = r * r * sqrt(1.0 + f * f) * sin(th);
        return av ? th * v : v;
    }
};

__device__ __forceinline__ double psi_of(double th, double ph, double TH)
{
    return acos(cos(TH) * cos(th) + sin(TH) * sin(th) * cos(ph));
}

__device__ __forceinline__ double phi_of(double th, double TH, double psi)
{
    double cp = cos(psi) - cos(TH) * cos(th);
    cp /= sin(TH) * sin(th);
    if (!(-1.0 <= cp && cp <= 1.0)) return -1.0;
    return acos(cp);
}

// phi_of with the sines and cosines already formed (cos psi, cos TH, sin TH
// and those of th): the same operations on the same values
__device__ __forceinline__ double phi_of_cs(double cth, double sth, double cTH, double sTH, double cpsi)
{
    double cp = cpsi - cTH * cth;
    cp /= sTH * sth;
    if (!(-1.0 <= cp && cp <= 1.0)) return -1.0;
    return acos(cp);
}

__device__ __forceinline__ double span_in(double a, double b, double LB, double UB)
{
    int ac = (LB <= a && a <= UB), bc = (LB <= b && b <= UB);
    if (!ac) ac = 2 * (a > UB);
    if (!bc) bc = 2 * (b > UB);
    if (ac == bc) return ac == 1 ? b - a : 0.0;
    if (bc == 1 && ac == 0) return b - LB;
    if (bc == 2 && ac == 1) return UB - a;
    if (bc == 2 && ac == 0) return UB - LB;
    return 0.0;
}

// azimuthal extent of (ceding region minus superseding region) at colatitude
// th, clipped to the cell [pa, pb] when `cell` is set.  prep() forms the
// circles' sines and cosines once per integral instead of at every node.
struct RegionFn {
    double eps, zeta, colat, cedeR, superR, superAzi, superColat, pa, pb, R_eq, Omega;
    int shape, lorentz, cell;
    double cCo = 0.0, sCo = 0.0, cCR = 0.0, cSu = 0.0, sSu = 0.0, cSR = 0.0;
    __device__ void prep()
    {
        cCo = cos(colat); sCo = sin(colat); cCR = cos(cedeR);
        cSu = cos(superColat); sSu = sin(superColat); cSR = cos(superR);
    }
    __device__ double span(double L, double U) const { return cell ? span_in(pa, pb, L, U) : U - L; }
    __device__ double operator()(double th) const
    {
        if (eq12(th, 0.0)) return 0.0;
        double mu = cos(th);
        const double sth = sin(th);
        double r = radius_norm(mu, eps, zeta, shape);
        double f = f_theta(mu, r, eps, zeta, shape);
        double aLB, aUB, cLB, cUB, dphi = 0.0;
        aUB = phi_of_cs(mu, sth, cCo, sCo, cCR);
        if (eq12(aUB, -1.0)) {
            if (th + colat < cedeR) { aUB = PI; aLB = -PI; }
            else return 0.0;
        } else aLB = -aUB;
        if (fabs(th - superColat) > superR) {
            cLB = cUB = 0.0;
        } else {
            cUB = phi_of_cs(mu, sth, cSu, sSu, cSR);
            if (eq12(cUB, -1.0)) return 0.0;
            cLB = -cUB;
            cUB += superAzi;
            cLB += superAzi;
            if (cUB > PI) cUB -= TWO_PI;
            if (cLB < -PI) cLB += TWO_PI;
        }
        if (cUB >= cLB) {
            if (cLB >= aUB || cUB <= aLB) dphi = span(aLB, aUB);
            else if (cLB >= aLB && cUB <= aUB) dphi = span(aLB, cLB) + span(cUB, aUB);
            else if (cLB < aLB && aLB < cUB && cUB <= aUB) dphi = span(cUB, aUB);
            else if (aLB <= cLB && cLB < aUB && aUB < cUB) dphi = span(aLB, cLB);
            else if (cLB < aLB && cUB > aUB) return 0.0;
        } else {
            if (cUB >= aUB || cLB <= aLB) return 0.0;
            else if (cUB >= aLB && cLB <= aUB) dphi = span(cUB, cLB);
            else if (cUB < aLB && aLB < cLB && cLB <= aUB) dphi = span(aLB, cLB);
            else if (aLB <= cUB && cUB < aUB && aUB < cLB) dphi = span(cUB, aUB);
            else if (cUB < aLB && cLB > aUB) dphi = span(aLB, aUB);
        }
        double w = 1.0;
        if (!cell && lorentz) {
            double rsr = 2.0 * zeta / r;
            double beta = R_eq * r * Omega * sth / (C_SI * sqrt(1.0 - rsr));
            w = 1.0 / sqrt(1.0 - beta * beta);
        }
        return w * dphi * r * r * sqrt(1.0 + f * f) * sth;
    }
};

__device__ __noinline__ double wave_area(double lo, double hi, double R_eq, double eps, double zeta, int shape, int av)
{
    AreaFn f{eps, zeta, av, shape};
    double r, e;
    wave_qag61(f, lo, hi, 0.0, MESH_EPSREL, QAG_LIMIT, r, e);
    return R_eq * R_eq * r;
}

__device__ double wave_spot(double ta, double tb, double R_eq, double eps, double zeta, int shape,
                            double colat, double cedeR, double superR, double superColat,
                            double superAzi, double Omega)
{
    RegionFn f{eps, zeta, colat, cedeR, superR, superAzi, superColat, 0.0, 0.0, R_eq, Omega,
               shape, 0, 0};
    f.prep();
    double r, e;
    wave_qag61(f, ta, tb, 0.0, MESH_EPSREL, QAG_LIMIT, r, e);
    return R_eq * R_eq * r;
}

__device__ double theta_given_phi(double ph, double colat, double rad, double lo, double hi)
{
    double x = sin(colat) * sin(ph) / sin(rad);
    if (!(x <= 1.0)) return -1.0;
    x = asin(x);
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) x = PI - x;
        double th = tan(0.5 * (colat - rad)) / sin(0.5 * (x - ph));
        th *= sin(0.5 * (x + ph));
        th = atan(th);
        th = th < 0.0 ? -2.0 * th : 2.0 * th;
        if (lo <= th && th <= hi) return th;
    }
    return -1.0;
}

__device__ double cede_edge_phi(double th, double colat, double cedeR, double pa, double pb)
{
    double v = phi_of(th, colat, cedeR);
    if (v >= 0.0 && !(pa <= v && v <= pb) && !(pa <= -v && -v <= pb)) v = -1.0;
    return v;
}

__device__ double super_edge_phi(double th, double sColat, double sR, double sAzi, double pa,
                                 double pb, bool strict_second)
{
    if (fabs(th - sColat) > sR) return -1.0;
    double v = phi_of(th, sColat, sR);
    if (v >= 0.0) {
        v += sAzi;
        if (v > PI) v -= TWO_PI;
        if (v < -PI) v += TWO_PI;
        if (!(pa <= v && v <= pb)) v = -1.0;
        else if (v < 0.0) v = -v;
    }
    if (v < 0.0) {
        v = phi_of(th, sColat, sR);
        bool take = strict_second ? (v > 0.0) : (v >= 0.0);
        if (take) {
            v = -v + sAzi;
            if (v > PI) v -= TWO_PI;
            if (v < -PI) v += TWO_PI;
            if (!(pa <= v && v <= pb)) v = -1.0;
            else if (v < 0.0) v = -v;
        }
    }
    return v;
}

__device__ __forceinline__ double take_first(double& c0, double& c1, double& c2, double& c3,
                                             double ta, double tb)
{
    if (c0 >= 0.0) { double v = c0; c0 = -1.0; return v; }
    if (c1 >= 0.0) { double v = c1; c1 = -1.0; return v; }
    if (c2 >= 0.0) { c2 = -1.0; return ta; }
    if (c3 >= 0.0) { c3 = -1.0; return tb; }
    return -1.0;
}

// colatitudes where the circle of angular radius rad about (colatitude TH,
// azimuth 0) meets the meridian at azimuth ph: cos(rad) = cos TH cos t +
// sin TH sin t cos ph, i.e. A cos t + B sin t = C
__device__ __forceinline__ void circle_meridian(double TH, double rad, double ph, double* t)
{
    const double A = cos(TH), B = sin(TH) * cos(ph), C = cos(rad);
    const double Rn = sqrt(A * A + B * B);
    t[0] = t[1] = -1.0;
    if (Rn == 0.0 || fabs(C) > Rn) return;
    const double d = atan2(B, A), w = acos(C / Rn);
    t[0] = d - w;
    t[1] = d + w;
}

// Sum of QAG-GK61 integrals over [lo, hi] split at the sorted points bp in
// (lo, hi): the cell integrand's kinks (where a region boundary crosses a cell
// edge) and square-root ends (a circle's extreme colatitudes) become piece
// ends, where the rule converges in a pass or two instead of bisecting toward
// them.  Any split is exact; a missed point only costs iterations.
template <class Fn>
__device__ double wave_pieces(const Fn& f, double lo, double hi, const double* bp, int nbp)
{
    double total = 0.0, a = lo;
    for (int k = 0; k <= nbp; k++) {
        double b = hi;
        if (k < nbp) {
            if (!(bp[k] > a && bp[k] < hi)) continue;
            b = bp[k];
        }
        double r, e;
        wave_qag61(f, a, b, 0.0, MESH_EPSREL, QAG_LIMIT, r, e);
        total += r;
        a = b;
    }
    return total;
}

// mesh_tools.pyx:429-725 (wave-cooperative).  bp: 16 doubles of LDS owned by
// the calling wavefront (its lanes hold the same breakpoints; in LDS the
// sort and the piece loop index them without scratch memory)
__device__ __noinline__ double wave_cell(double ta, double tb, double pa, double pb, double R_eq, double eps,
                            double zeta, int shape, double colat, double cedeR, double superR,
                            double superAzi, double superColat, double* bp)
{
    RegionFn f{eps, zeta, colat, cedeR, superR, superAzi, superColat, pa, pb, R_eq, 0.0, shape, 0, 1};
    f.prep();
    // candidate breakpoints, sorted
    int nbp = 0;
    auto add = [&](double t) { if (t > ta && t < tb && nbp < 16) bp[nbp++] = t; };
    double t2[2];
    circle_meridian(colat, cedeR, pa, t2); add(t2[0]); add(t2[1]);
    circle_meridian(colat, cedeR, pb, t2); add(t2[0]); add(t2[1]);
    add(colat - cedeR); add(colat + cedeR); add(cedeR - colat);
    if (superR > 0.0) {
        circle_meridian(superColat, superR, pa - superAzi, t2); add(t2[0]); add(t2[1]);
        circle_meridian(superColat, superR, pb - superAzi, t2); add(t2[0]); add(t2[1]);
        add(superColat - superR); add(superColat + superR); add(superR - superColat);
    }
    for (int q = 1; q < nbp; q++) {
        double v = bp[q];
        int r = q - 1;
        while (r >= 0 && bp[r] > v) { bp[r + 1] = bp[r]; r--; }
        bp[r + 1] = v;
    }
    double integral = wave_pieces(f, ta, tb, bp, nbp), e;
    if (eq12(integral, 0.0)) {
        double tl0 = theta_given_phi(pa, colat, cedeR, ta, tb);
        double tl1 = theta_given_phi(pa - superAzi, superColat, superR, ta, tb);
        double tr0 = theta_given_phi(pb, colat, cedeR, ta, tb);
        double tr1 = theta_given_phi(pb - superAzi, superColat, superR, ta, tb);
        double top0 = cede_edge_phi(ta, colat, cedeR, pa, pb);
        double top1 = super_edge_phi(ta, superColat, superR, superAzi, pa, pb, true);
        double bot0 = cede_edge_phi(tb, colat, cedeR, pa, pb);
        double bot1 = super_edge_phi(tb, superColat, superR, superAzi, pa, pb, false);
        double t0 = take_first(tl0, tr0, top0, bot0, ta, tb);
        double t1 = take_first(tl0, tr0, top0, bot0, ta, tb);
        double t2 = take_first(tl1, tr1, top1, bot1, ta, tb);
        double t3 = take_first(tl1, tr1, top1, bot1, ta, tb);
        double tmin = tb, tmax = ta;
        double ts[4] = {t0, t1, t2, t3};
        for (int q = 0; q < 4; q++) {
            if (ts[q] >= 0.0) {
                if (ts[q] < tmin) tmin = ts[q];
                if (ts[q] > tmax) tmax = ts[q];
            }
        }
        if (tmin < tmax) integral = wave_pieces(f, tmin, tmax, bp, nbp);
    }
    (void)e;
    return R_eq * R_eq * integral;
}

__device__ __forceinline__ double lin(double start, double stop, int num, int k)
{
    if (num == 1) return start;
    if (k == num - 1) return stop;
    double step = (stop - start) / (double)(num - 1);
    return (double)k * step + start;
}

__device__ void clamp_even(double& s, int lo, int hi)
{
    if (s < (double)lo) s = (double)lo;
    else if (s > (double)hi) s = (double)hi;
    if (((long)s) % 2 != 0) s += 1.0;
}

// ---------------------------------------------------------------- kernels

// One block of 4 waves per hot region: the up-to-four allocation integrals
// run concurrently, then lane 0 of wave 0 sizes both members.
__global__ void __launch_bounds__(256) k_allocate(const RegionAlloc* __restrict__ regions,
                                                  Member* __restrict__ members, int S_max)
{
    const RegionAlloc R = regions[blockIdx.x];
    const int wave = threadIdx.x >> 6;
    __shared__ double vals[4];
    __shared__ double lims_s[2], lims_c[2], cellfac[2];
    // super region limits (possibly reflected through the equator)
    double sColat = R.superColat, hColat = R.holeColat, s0, s1, sfac;
    if (R.superColat - R.superR < 0.0 || R.superColat + R.superR > PI) {
        s0 = 0.0;
        if (R.superColat + R.superR > PI) {
            s1 = PI - R.superColat + R.superR;
            sColat = PI - R.superColat;
            hColat = PI - R.holeColat;
        } else s1 = R.superColat + R.superR;
        sfac = TWO_PI;
    } else {
        s0 = R.superColat - R.superR;
        s1 = R.superColat + R.superR;
        sfac = 2.0 * asin(sin(R.superR) / sin(R.superColat));
    }
    bool cede = !eq12(R.cedeR, 0.0);
    double cColat = R.cedeColat, sColat2 = R.superColat, c0 = 0.0, c1 = 0.0, cfac = 0.0;
    if (cede) {
        if (R.cedeColat - R.cedeR < 0.0 || R.cedeColat + R.cedeR > PI) {
            c0 = 0.0;
            if (R.cedeColat + R.cedeR > PI) {
                c1 = PI - R.cedeColat + R.cedeR;
                cColat = PI - R.cedeColat;
                sColat2 = PI - R.superColat;
            } else c1 = R.cedeColat + R.cedeR;
            cfac = TWO_PI;
        } else {
            c0 = R.cedeColat - R.cedeR;
            c1 = R.cedeColat + R.cedeR;
            cfac = 2.0 * asin(sin(R.cedeR) / sin(R.cedeColat));
        }
    }
    double v = 0.0;
    if (wave == 0) v = wave_area(s0, s1, R.R_eq, R.epsilon, R.zeta, R.shape, 0);
    else if (wave == 1)
        v = wave_spot(s0, s1, R.R_eq, R.epsilon, R.zeta, R.shape, sColat, R.superR, R.holeR, hColat,
                      R.holeAzi, R.Omega);
    else if (wave == 2 && cede) v = wave_area(c0, c1, R.R_eq, R.epsilon, R.zeta, R.shape, 0);
    else if (wave == 3 && cede)
        v = wave_spot(c0, c1, R.R_eq, R.epsilon, R.zeta, R.shape, cColat, R.cedeR, R.superR, sColat2,
                      R.superAzi, R.Omega);
    if ((threadIdx.x & 63) == 0) vals[wave] = v;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double s_cell = sfac * vals[0], s_area = vals[1];
    double s_sqrt, s_num = 0.0, c_sqrt = 0.0, c_num = 0.0;
    if (!cede) {
        if (eq12(s_area, 0.0)) s_area = s_cell / 1000.0;
        s_sqrt = ceil(sqrt((double)R.num_cells * s_cell / s_area));
        clamp_even(s_sqrt, R.min_sqrt, R.max_sqrt);
        s_num = s_sqrt * s_sqrt;
    } else {
        double c_cell = cfac * vals[2], c_area = vals[3];
        if (eq12(s_area, 0.0)) s_area = s_cell / 1000.0;
        if (eq12(c_area, 0.0)) c_area = c_cell / 1000.0;
        double y = c_area / s_area - 1.0, sn, cn;
        if (eq12(y, 0.0)) { sn = 0.5 * R.num_cells; cn = 0.5 * R.num_cells; }
        else { sn = (double)R.num_cells * ((sqrt(1.0 + y) - 1.0) / y); cn = (double)R.num_cells - sn; }
        s_sqrt = ceil(sqrt(sn * s_cell / s_area));
        c_sqrt = ceil(sqrt(cn * c_cell / c_area));
        clamp_even(s_sqrt, R.min_sqrt, R.max_sqrt);
        clamp_even(c_sqrt, R.min_sqrt, R.max_sqrt);
        s_num = s_sqrt * s_sqrt;
        c_num = c_sqrt * c_sqrt;
    }
    Member* m0 = members + R.member0;
    m0->S = (int)s_sqrt;
    m0->numCell = s_num;
    if (m0->S > S_max) m0->status = 3;
    if (R.member1 >= 0) {
        Member* m1 = members + R.member1;
        m1->S = (int)c_sqrt;
        m1->numCell = c_num;
        if (m1->S > S_max) m1->status = 3;
    }
}

__device__ __forceinline__ void member_lims(const Member& m, double& lo, double& hi, double& colat,
                                            double& superColat, int& invert, double& dphi,
                                            double& bphi)
{
    colat = m.cedeColat;
    superColat = m.superColat;
    invert = 0;
    bphi = 0.0;
    if (m.polar) {
        lo = 0.0;
        if (m.cedeColat + m.cedeR > PI) {
            invert = 1;
            hi = PI - m.cedeColat + m.cedeR;
            colat = PI - m.cedeColat;
            superColat = PI - m.superColat;
        } else hi = m.cedeColat + m.cedeR;
        dphi = TWO_PI / (double)m.S;
    } else {
        lo = m.cedeColat - m.cedeR;
        hi = m.cedeColat + m.cedeR;
        bphi = asin(sin(m.cedeR) / sin(m.cedeColat));
        dphi = 2.0 * bphi / (double)m.S;
    }
}

// Area table (mesh.pyx:79-115): entry q is the area from table colatitude q
// (linspace(hi, lo, table_n)) to the far boundary hi, and the total is the
// area over [lo, hi].  The reference integrates each entry separately; here
// entry q is the running sum of the integrals over the table's intervals
// [ct_q, ct_(q-1)], each under (hi - lo) / 999 wide, where the smooth area
// integrand makes a 7-point Gauss-Legendre rule exact to rounding.  One block
// per member: a lane per interval, then a block-wide inclusive scan.
__global__ void __launch_bounds__(1024) k_mesh_table(const Member* __restrict__ members,
                                                     double* __restrict__ area_tab,
                                                     double* __restrict__ total, int table_n)
{
    // Gauss-Legendre, 7 points on [-1, 1] (abscissae and weights to 16 digits)
    constexpr double GX[4] = {0.0, 0.4058451513773971669, 0.7415311855993944399, 0.9491079123427585245};
    constexpr double GW[4] = {0.4179591836734693878, 0.3818300505051189449, 0.2797053914892766679,
                              0.1294849661688696933};
    const int mi = blockIdx.x, tid = threadIdx.x;
    const Member m = members[mi];
    if (!m.active || m.status) return;
    double lo, hi, colat, sc, dphi, bphi;
    int inv;
    member_lims(m, lo, hi, colat, sc, inv, dphi, bphi);
    AreaFn f{m.epsilon, m.zeta, 0, m.shape};
    __shared__ double part[1024];
    double v = 0.0;
    if (tid >= 1 && tid < table_n) {
        const double a = lin(hi, lo, table_n, tid), b = lin(hi, lo, table_n, tid - 1);
        const double c = 0.5 * (a + b), h = 0.5 * (b - a);
        double sum = GW[0] * f(c);
#pragma unroll
        for (int k = 1; k < 4; k++) sum += GW[k] * (f(c - h * GX[k]) + f(c + h * GX[k]));
        v = sum * h;
    }
    part[tid] = v;
    __syncthreads();
    // inclusive scan (Hillis-Steele)
    for (int o = 1; o < 1024; o <<= 1) {
        const double add = tid >= o ? part[tid - o] : 0.0;
        __syncthreads();
        part[tid] += add;
        __syncthreads();
    }
    const double R2 = m.R_eq * m.R_eq;   // as wave_area scales its integral
    if (tid < table_n) area_tab[(size_t)mi * table_n + tid] = R2 * part[tid];
    if (tid == 0) total[mi] = R2 * part[table_n - 1];
}

// The cell mesh of a member, in three launches:
//   k_ring_bounds   one wavefront per parallel: its bounding colatitudes from
//                   the equal-area table (Steffen inversion) and its centroid
//                   colatitude (an area-weighted integral), and the per-ring
//                   outputs;
//   k_mesh_classify one wavefront per parallel, lanes over its cells: each
//                   cell's kind from its corners and centre (mesh.pyx's rule),
//                   full and empty cells written, boundary cells appended to
//                   a list;
//   k_mesh_cells    wavefronts in a grid-stride loop over the list: each
//                   boundary cell's area by the split QAG integral.
// Splitting the parallels' set-up from the cells keeps every wavefront busy
// (the set-up is serial FP64 work of one wavefront per parallel), and the list
// spreads the boundary cells, a few per parallel on most and a whole row on
// the parallels at the regions' top and bottom, evenly over the chip.

// (lower, upper, centroid) colatitudes of each parallel: ringlim[(member * S_max + ring) * 3 + k]
__global__ void __launch_bounds__(256) k_ring_bounds(const Member* __restrict__ members,
                                                     const double* __restrict__ area_tab,
                                                     const double* __restrict__ total, int table_n, Layout L,
                                                     double* __restrict__ theta, double* __restrict__ rr,
                                                     double* __restrict__ maxAlpha, double* __restrict__ cos_gamma,
                                                     double* __restrict__ effGrav, double* __restrict__ ringlim)
{
    const int mi = blockIdx.y, tid = threadIdx.x, lane = tid & 63;
    const int i = blockIdx.x * 4 + (tid >> 6);
    const Member m = members[mi];
    if (!m.active || m.status || blockIdx.x * 4 >= m.S) return;
    const int S = m.S;
    double lo, hi, colat, sColat, dphi, bphi;
    int inv;
    member_lims(m, lo, hi, colat, sColat, inv, dphi, bphi);
    double cell = (m.polar ? TWO_PI : 2.0 * bphi) * total[mi];
    cell /= m.numCell;
    const double eta = cell / dphi;
    // the member's area table and its colatitudes in LDS: the Steffen
    // inversions' interval searches read them about ten times in a row
    __shared__ double ct[1024], at[1024];
    const double* atg = area_tab + (size_t)mi * table_n;
    for (int k = tid; k < table_n; k += 256) {
        ct[k] = lin(hi, lo, table_n, k);
        at[k] = atg[k];
    }
    __syncthreads();
    if (i >= S) return;
    const size_t ro = (size_t)mi * L.S_max;
    // the reference subtracts eta ring by ring
    double i_eta = (double)S * eta, e_lo = 0.0, e_hi = 0.0;
    for (int k = 0; k <= i && k < S - 1; k++) {
        i_eta -= eta;
        if (k == i - 1) e_lo = i_eta;
        if (k == i) e_hi = i_eta;
    }
    const double lower = (i == 0) ? lo : steffen_eval(at, ct, table_n, e_lo, nullptr);
    const double upper = (i == S - 1) ? hi : steffen_eval(at, ct, table_n, e_hi, nullptr);
    const double cc = wave_area(lower, upper, m.R_eq, m.epsilon, m.zeta, m.shape, 1) / eta;
    if (lane == 0) {
        double* rl = ringlim + (ro + i) * 3;
        rl[0] = lower; rl[1] = upper; rl[2] = cc;
        double mu = cos(cc);
        double rn = radius_norm(mu, m.epsilon, m.zeta, m.shape);
        double fth = f_theta(mu, rn, m.epsilon, m.zeta, m.shape);
        double cg = 1.0 / sqrt(1.0 + fth * fth);
        rr[ro + i] = rn * m.R_eq;
        effGrav[ro + i] = effective_gravity(mu, m.R_eq, m.zeta, m.epsilon, m.shape);
        cos_gamma[ro + i] = cg;
        maxAlpha[ro + i] = HALF_PI + acos(cg);
        theta[ro + i] = inv ? PI - cc : cc;
    }
}

// psi_of(th, ph, TH) <= R as the cosine of psi against cos R: the same
// argument psi_of forms, and its NaN (an argument beyond [-1, 1]) counts as
// outside, as the comparison with a NaN angle does
__device__ __forceinline__ bool within(double cTH, double sTH, double cth, double sth, double cph, double cR)
{
    const double arg = cTH * cth + sTH * sth * cph;
    return arg <= 1.0 && arg >= cR;
}
__device__ __forceinline__ bool beyond(double cTH, double sTH, double cth, double sth, double cph, double cR)
{
    const double arg = cTH * cth + sTH * sth * cph;
    return arg >= -1.0 && arg <= 1.0 && arg <= cR;
}

// azimuth of the left edge of column j: the reference accumulates
// left += delta_phi cell by cell
__device__ __forceinline__ double left_edge(double first_left, double dphi, int j)
{
    double left = first_left;
    for (int q = 0; q < j; q++) left = left + dphi;
    return left;
}

// one wavefront per parallel, lanes over its cells, eight parallels of one
// member per block; blist entries (member, ring << 16 | column).  The
// boundary cells of a block's round of 64 columns go onto the list with one
// atomic for the block (the list's counter is shared by every block, and an
// atomic per wavefront serialised there).
constexpr int CLASSIFY_WAVES = 8;
__global__ void __launch_bounds__(64 * CLASSIFY_WAVES) k_mesh_classify(const Member* __restrict__ members,
                                                       const double* __restrict__ total, Layout L,
                                                       const double* __restrict__ ringlim,
                                                       double* __restrict__ phi, double* __restrict__ cellArea,
                                                       int2* __restrict__ blist, int* __restrict__ bcount)
{
    const int mi = blockIdx.y, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int i = blockIdx.x * CLASSIFY_WAVES + w;
    const Member m = members[mi];
    if (!m.active || m.status) return;   // the whole block
    const int S = m.S;
    if (blockIdx.x * CLASSIFY_WAVES >= S) return;   // the whole block
    const bool ring = i < S;                         // this wavefront's parallel exists
    __shared__ int cnt[CLASSIFY_WAVES], off[CLASSIFY_WAVES], base_s;
    double lo, hi, colat, sColat, dphi, bphi;
    int inv;
    member_lims(m, lo, hi, colat, sColat, inv, dphi, bphi);
    double cell = (m.polar ? TWO_PI : 2.0 * bphi) * total[mi];
    cell /= m.numCell;
    const size_t ro = (size_t)mi * L.S_max, co = (size_t)mi * L.S_max * L.S_max;
    const double first_left = m.polar ? -PI : -bphi;
    if (i == 0) {
        for (int j = lane; j < S; j += 64) {
            double p0 = m.polar ? lin(-PI + 0.5 * dphi, PI - 0.5 * dphi, S, j)
                                : lin(-bphi + 0.5 * dphi, bphi - 0.5 * dphi, S, j);
            double v = p0 + m.phi_offset;
            if (m.antiphased) v += PI;
            phi[ro + j] = v;
        }
    }
    const double* rl = ringlim + (ro + (ring ? i : 0)) * 3;   // a row that exists
    const double lower = rl[0], upper = rl[1], cc = rl[2];
    const double sR = m.superR, sAzi = m.superAzi, cR = m.cedeR;
    double s1a = 0, s1b = 0, s2a = 0, s2b = 0, s3a = 0, s3b = 0, s4a = 0, s4b = 0;
    if (sR > 0.0) {
        if (sColat - sR < 0.0) {
            s1a = sR - sColat; s1b = sAzi + PI; if (s1b > PI) s1b -= TWO_PI;
            s2a = sColat + sR; s2b = sAzi;
        } else if (sColat + sR > PI) {
            s1a = sColat - sR; s1b = sAzi;
            s2a = TWO_PI - sColat - sR; s2b = sAzi + PI; if (s2b > PI) s2b -= TWO_PI;
        } else {
            s1a = sColat - sR; s1b = sAzi;
            s2a = acos(cos(sColat) / cos(sR));
            s2b = sAzi + asin(sin(sR) / sin(sColat)); if (s2b > PI) s2b -= TWO_PI;
            s3a = sColat + sR; s3b = sAzi;
            s4a = s2a;
            s4b = sAzi - asin(sin(sR) / sin(sColat)); if (s4b < -PI) s4b += TWO_PI;
        }
    }
    const bool mirror = eq12(sColat, colat) && eq12(sAzi, 0.0);
    const int j_max = mirror ? S / 2 : S;
    // the parallel's and the circles' sines and cosines, once
    const double cLo = cos(lower), sLo = sin(lower), cUp = cos(upper), sUp = sin(upper);
    const double cCc = cos(cc), sCc = sin(cc);
    const double cCo = cos(colat), sCo = sin(colat), cSu = cos(sColat), sSu = sin(sColat);
    const double ccR = cos(cR), csR = cos(sR);
    double* ca = cellArea + co + (size_t)(ring ? i : 0) * S;
    // j_max is the member's, so every wavefront of the block takes the same rounds
    for (int j0 = 0; j0 < j_max; j0 += 64) {
        const int j = j0 + lane;
        int kind = 0;
        if (ring && j < j_max) {
            const double left = left_edge(first_left, dphi, j);
            const double right = left + dphi, mid = left + 0.5 * dphi;
            const double cl = cos(left), cr = cos(right), cm = cos(mid);
            const double cls = cos(left - sAzi), crs = cos(right - sAzi), cms = cos(mid - sAzi);
            int p1, p2, p3, p4, p5;
            if (m.polar && i == 0) {
                p1 = p3 = (colat <= cR);
                if (p1 == 1) { p1 = (sR <= sColat); p3 = p1; } else p1 = p3 = 2;
            } else {
                p1 = within(cCo, sCo, cLo, sLo, cl, ccR);
                p3 = within(cCo, sCo, cLo, sLo, cr, ccR);
                p1 = p1 == 1 ? beyond(cSu, sSu, cLo, sLo, cls, csR) : 2;
                p3 = p3 == 1 ? beyond(cSu, sSu, cLo, sLo, crs, csR) : 2;
            }
            p2 = within(cCo, sCo, cUp, sUp, cl, ccR);
            p4 = within(cCo, sCo, cUp, sUp, cr, ccR);
            p5 = within(cCo, sCo, cCc, sCc, cm, ccR);
            p2 = p2 == 1 ? beyond(cSu, sSu, cUp, sUp, cls, csR) : 2;
            p4 = p4 == 1 ? beyond(cSu, sSu, cUp, sUp, crs, csR) : 2;
            p5 = p5 == 1 ? beyond(cSu, sSu, cCc, sCc, cms, csR) : 2;
            if (!(p1 == p2 && p2 == p3 && p3 == p4 && p4 == p5)) kind = 2;
            else if (sR > 0.0) {
                auto inside = [&](double a, double b) {
                    return (lower <= a && a <= upper) && (left <= b && b <= right);
                };
                if (inside(s1a, s1b) || inside(s2a, s2b) || inside(s3a, s3b) || inside(s4a, s4b)) kind = 2;
                else if (p5 == 1) kind = 1;
            } else if (p5 == 1) kind = 1;
            const double v = kind == 1 ? cell : 0.0;
            ca[j] = v;
            if (mirror) ca[S - 1 - j] = v;
        }
        // boundary cells onto the list: the block's count in one atomic, each
        // wavefront's entries after those of the wavefronts before it
        const unsigned long long bm = __ballot(kind == 2);
        if (lane == 0) cnt[w] = __popcll(bm);
        __syncthreads();
        if (tid == 0) {
            int tot = 0;
            for (int q = 0; q < CLASSIFY_WAVES; q++) { off[q] = tot; tot += cnt[q]; }
            base_s = tot ? atomicAdd(bcount, tot) : 0;
        }
        __syncthreads();
        if (kind == 2) blist[base_s + off[w] + __popcll(bm & ((1ull << lane) - 1ull))] = make_int2(mi, (i << 16) | j);
        __syncthreads();   // cnt, off and base_s are rewritten by the next round
    }
}

// wavefronts over the boundary-cell list (grid-stride): each cell's area
__global__ void __launch_bounds__(256, 3) k_mesh_cells(const Member* __restrict__ members, const double* __restrict__ total,
                                                    Layout L, const double* __restrict__ ringlim,
                                                    double* __restrict__ cellArea, const int2* __restrict__ blist,
                                                    const int* __restrict__ bcount)
{
    const int lane = threadIdx.x & 63;
    const int nw = gridDim.x * 4;
    const int n = *bcount;
    __shared__ double bps[4][16];   // each wavefront's breakpoint list
    double* bp = bps[threadIdx.x >> 6];
    for (int q = blockIdx.x * 4 + (threadIdx.x >> 6); q < n; q += nw) {
        const int2 e = blist[q];
        const int mi = e.x, i = e.y >> 16, j = e.y & 0xffff;
        const Member m = members[mi];
        const int S = m.S;
        double lo, hi, colat, sColat, dphi, bphi;
        int inv;
        member_lims(m, lo, hi, colat, sColat, inv, dphi, bphi);
        const double* rl = ringlim + ((size_t)mi * L.S_max + i) * 3;
        const double first_left = m.polar ? -PI : -bphi;
        const double left = left_edge(first_left, dphi, j), right = left + dphi;
        const double v = wave_cell(rl[0], rl[1], left, right, m.R_eq, m.epsilon, m.zeta, m.shape, colat, m.cedeR,
                                   m.superR, m.superAzi, sColat, bp);
        const bool mirror = eq12(sColat, colat) && eq12(m.superAzi, 0.0);
        if (lane == 0) {
            double* ca = cellArea + (size_t)mi * L.S_max * L.S_max + (size_t)i * S;
            ca[j] = v;
            if (mirror) ca[S - 1 - j] = v;
        }
    }
}

}  // namespace xh

namespace xh {

// per-ring quantities the integrator needs, from the mesh of each member
__global__ void __launch_bounds__(256) k_ring_finalize(const Member* __restrict__ members, Layout L,
                                                       const double* __restrict__ rr,
                                                       const double* __restrict__ cellArea,
                                                       double* __restrict__ u, int* __restrict__ ring_ok,
                                                       double* __restrict__ logT)
{
    const int mi = blockIdx.x;
    const Member m = members[mi];
    if (!m.active || m.status) return;
    for (int i = threadIdx.x; i < m.S; i += blockDim.x) {
        size_t ro = (size_t)mi * L.S_max + i;
        u[ro] = m.r_s / rr[ro];
        const double* ca = cellArea + (size_t)mi * L.S_max * L.S_max + (size_t)i * m.S;
        int ok = 0;
        for (int j = 0; j < m.S; j++) if (ca[j] > 0.0) { ok = 1; break; }
        ring_ok[ro] = ok;
        logT[ro] = m.logT;
    }
}

// drop-in path: per-ring inputs from the reference's [S, S] arrays
__global__ void __launch_bounds__(256) k_ring_prep(int S, const double* __restrict__ theta2d,
                                                   const double* __restrict__ params,
                                                   const int* __restrict__ radiates,
                                                   double* __restrict__ theta, int* __restrict__ ring_ok,
                                                   double* __restrict__ logT, double* __restrict__ logg)
{
    for (int i = threadIdx.x; i < S; i += blockDim.x) {
        int J = -1;
        for (int j = 0; j < S; j++) if (radiates[(size_t)i * S + j] == 1) { J = j; break; }
        ring_ok[i] = J >= 0;
        logT[i] = J >= 0 ? params[((size_t)i * S + J) * 2] : 0.0;
        logg[i] = J >= 0 ? params[((size_t)i * S + J) * 2 + 1] : 0.0;
        theta[i] = theta2d[(size_t)i * S];
    }
}

// ---------------------------------------------------------------- closed mesh
// Equal-area mesh of the whole star for the time-invariant "elsewhere"
// component (xpsi/cellmesh/global_mesh.pyx:21-147).  k_closed_table: one
// wave per point of the 1000-point area table (area from a colatitude to the
// equator), plus the hemisphere area (point 0).
__global__ void __launch_bounds__(256) k_closed_table(const Member* __restrict__ members,
                                                      double* __restrict__ area_tab,
                                                      double* __restrict__ total, int table_n)
{
    const Member m = members[blockIdx.y];
    if (!m.active || m.status) return;
    int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q > table_n) return;
    double v;
    if (q == 0) v = wave_area(0.0, HALF_PI, m.R_eq, m.epsilon, m.zeta, m.shape, 0);
    else v = wave_area(lin(HALF_PI, 0.0, table_n, q - 1), HALF_PI, m.R_eq, m.epsilon, m.zeta, m.shape, 0);
    if ((threadIdx.x & 63) == 0) {
        if (q == 0) total[blockIdx.y] = v;
        else area_tab[(size_t)blockIdx.y * table_n + q - 1] = v;
    }
}

// One wave per northern ring: its bounding parallels by inverting the area
// table (Steffen), the area-weighted mean colatitude between them, and the
// ring's radius, surface tilt, gravity and maximum emission angle; the
// southern ring S-1-i mirrors it.  cellArea[member] = the common cell area.
__global__ void __launch_bounds__(64) k_closed_rings(const Member* __restrict__ members,
                                                     const double* __restrict__ area_tab,
                                                     const double* __restrict__ total, int table_n,
                                                     int S_max, double* __restrict__ colat,
                                                     double* __restrict__ rr, double* __restrict__ maxAlpha,
                                                     double* __restrict__ cos_gamma,
                                                     double* __restrict__ effGrav,
                                                     double* __restrict__ cellArea)
{
    const int mi = blockIdx.y, i = blockIdx.x, lane = threadIdx.x;
    const Member m = members[mi];
    if (!m.active || m.status) return;
    const int S = m.S, ncc = S / 2;
    if (i >= ncc) return;
    // the area table and its colatitudes in LDS for the interval searches
    __shared__ double ct[1024], at[1024];
    const double* atg = area_tab + (size_t)mi * table_n;
    for (int k = lane; k < table_n; k += 64) {
        ct[k] = lin(HALF_PI, 0.0, table_n, k);
        at[k] = atg[k];
    }
    __syncthreads();
    double cell = TWO_PI * total[mi] / ((double)S * (double)S);
    cell *= 2.0;
    const double eta = cell * (double)S / TWO_PI;
    // parallels[k] at i_eta = ncc eta - (k + 1) eta, by repeated subtraction
    double i_eta = (double)ncc * eta, e_lo = 0.0, e_hi = 0.0;
    for (int k = 0; k <= i && k < ncc - 1; k++) {
        i_eta -= eta;
        if (k == i - 1) e_lo = i_eta;
        if (k == i) e_hi = i_eta;
    }
    const double lower = (i == 0) ? 0.0 : steffen_eval(at, ct, table_n, e_lo, nullptr);
    const double upper = (i == ncc - 1) ? HALF_PI : steffen_eval(at, ct, table_n, e_hi, nullptr);
    const double cc = wave_area(lower, upper, m.R_eq, m.epsilon, m.zeta, m.shape, 1) / eta;
    if (lane == 0) {
        const size_t ro = (size_t)mi * S_max;
        const double mu = cos(cc);
        const double rn = radius_norm(mu, m.epsilon, m.zeta, m.shape);
        const double fth = f_theta(mu, rn, m.epsilon, m.zeta, m.shape);
        const double cg = 1.0 / sqrt(1.0 + fth * fth);
        const double eg = effective_gravity(mu, m.R_eq, m.zeta, m.epsilon, m.shape);
        const int k = S - 1 - i;
        colat[ro + i] = cc;
        colat[ro + k] = PI - cc;
        rr[ro + i] = rn * m.R_eq;
        rr[ro + k] = rn * m.R_eq;
        cos_gamma[ro + i] = cg;
        cos_gamma[ro + k] = cg;
        effGrav[ro + i] = eg;
        effGrav[ro + k] = eg;
        maxAlpha[ro + i] = HALF_PI + acos(cg);
        maxAlpha[ro + k] = HALF_PI + acos(cg);
        if (i == 0) cellArea[mi] = cell;
    }
}

}  // namespace xh
