/*
 * Oracle: surface cell mesh of a hot region.  TEST INFRASTRUCTURE ONLY.
 * Restates xpsi/cellmesh/mesh_tools.pyx:18-1099 (shape, area and cell
 * integrals, allocate_cells), xpsi/cellmesh/mesh.pyx:23-417
 * (construct_spot_cellMesh), xpsi/cellmesh/polar_mesh.pyx:23-430
 * (construct_polar_cellMesh) and
 * xpsi/surface_radiation_field/effective_gravity_universal.pyx:10-71.
 * The reference's gsl_integration_cquad(epsrel 1e-8) calls are replaced by
 * QAG-GK61 at OR_MESH_EPSREL; boundary-cell integrals are split at their
 * kinks (cell_breakpoints) before the rule is applied.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "xpsi_oracle.h"

#define TWO_PI (2.0 * M_PI)
static const double C_SI = 2.99792458e8;

static int eq12(double x, double y) { return fabs(x - y) < 1.0e-12; }

/* mesh_tools.pyx:18-56 */
static double radius_norm(double mu, double eps, double zeta, int shape)
{
    return shape == 0 ? 1.0 + eps * (-0.788 + 1.030 * zeta) * mu * mu : 1.0;
}

/* mesh_tools.pyx:58-97 */
static double f_theta(double mu, double rn, double eps, double zeta, int shape)
{
    double d = -2.0 * eps * (-0.788 + 1.030 * zeta) * mu * sqrt(1.0 - mu * mu);
    return shape == 0 ? d / (rn * sqrt(1.0 - 2.0 * zeta / rn)) : 0.0;
}

/* effective_gravity_universal.pyx:10-71 */
double or_effective_gravity(double mu, double R_eq, double x, double eps,
                            int shape)
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

typedef struct { double eps, zeta; int av, shape; } area_par;

/* mesh_tools.pyx:99-120 */
static double area_integrand(double th, void *vp)
{
    area_par *p = (area_par *)vp;
    if (eq12(th, 0.0)) return 0.0;
    double mu = cos(th);
    double r = radius_norm(mu, p->eps, p->zeta, p->shape);
    double f = f_theta(mu, r, p->eps, p->zeta, p->shape);
    double v = r * r * sqrt(1.0 + f * f) * sin(th);
    return p->av ? th * v : v;
}

static double mesh_quad(or_fn f, void *p, double a, double b)
{
    double r, e;
    or_qag61(f, p, a, b, 0.0, OR_MESH_EPSREL, 100, &r, &e);
    return r;
}

/* mesh_tools.pyx:122-153 */
static double integrate_area(double lo, double hi, double R_eq, double eps,
                             double zeta, int shape, int av)
{
    area_par p = {eps, zeta, av, shape};
    return R_eq * R_eq * mesh_quad(area_integrand, &p, lo, hi);
}

/* mesh_tools.pyx:156-160 */
static double eval_psi(double th, double ph, double TH)
{
    return acos(cos(TH) * cos(th) + sin(TH) * sin(th) * cos(ph));
}

/* mesh_tools.pyx:265-274 */
static double eval_phi(double th, double TH, double psi)
{
    double cp = cos(psi) - cos(TH) * cos(th);
    cp /= sin(TH) * sin(th);
    if (!(-1.0 <= cp && cp <= 1.0)) return -1.0;
    return acos(cp);
}

/* mesh_tools.pyx:276-301 */
static double get_interval(double a, double b, double LB, double UB)
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

typedef struct {
    double eps, zeta, colat, cedeR, superR, superAzi, superColat;
    double phi_a, phi_b, R_eq, Omega;
    int shape, lorentz, cell;
} region_par;

/*
 * Azimuthal extent of the ceding region minus the superseding region at
 * colatitude th, either clipped to [phi_a, phi_b] (cell integrand,
 * mesh_tools.pyx:303-388) or unclipped (spot integrand, :728-813).
 */
static double region_integrand(double th, void *vp)
{
    region_par *p = (region_par *)vp;
    if (eq12(th, 0.0)) return 0.0;
    double mu = cos(th);
    double r = radius_norm(mu, p->eps, p->zeta, p->shape);
    double f = f_theta(mu, r, p->eps, p->zeta, p->shape);
    double aLB, aUB, cLB, cUB, dphi = 0.0;
    aUB = eval_phi(th, p->colat, p->cedeR);
    if (eq12(aUB, -1.0)) {
        if (th + p->colat < p->cedeR) { aUB = M_PI; aLB = -M_PI; }
        else return 0.0;
    } else aLB = -aUB;
    if (fabs(th - p->superColat) > p->superR) {
        cLB = cUB = 0.0;
    } else {
        cUB = eval_phi(th, p->superColat, p->superR);
        if (eq12(cUB, -1.0)) return 0.0;
        cLB = -cUB;
        cUB += p->superAzi;
        cLB += p->superAzi;
        if (cUB > M_PI) cUB -= TWO_PI;
        if (cLB < -M_PI) cLB += TWO_PI;
    }
#define SPAN(L, U) (p->cell ? get_interval(p->phi_a, p->phi_b, (L), (U)) : (U) - (L))
    if (cUB >= cLB) {
        if (cLB >= aUB || cUB <= aLB) dphi = SPAN(aLB, aUB);
        else if (cLB >= aLB && cUB <= aUB) dphi = SPAN(aLB, cLB) + SPAN(cUB, aUB);
        else if (cLB < aLB && aLB < cUB && cUB <= aUB) dphi = SPAN(cUB, aUB);
        else if (aLB <= cLB && cLB < aUB && aUB < cUB) dphi = SPAN(aLB, cLB);
        else if (cLB < aLB && cUB > aUB) return 0.0;
    } else {
        if (cUB >= aUB || cLB <= aLB) return 0.0;
        else if (cUB >= aLB && cLB <= aUB) dphi = SPAN(cUB, cLB);
        else if (cUB < aLB && aLB < cLB && cLB <= aUB) dphi = SPAN(aLB, cLB);
        else if (aLB <= cUB && cUB < aUB && aUB < cLB) dphi = SPAN(cUB, aUB);
        else if (cUB < aLB && cLB > aUB) dphi = SPAN(aLB, aUB);
    }
#undef SPAN
    double w = 1.0;
    if (!p->cell && p->lorentz) {
        double rsr = 2.0 * p->zeta / r;
        double beta = p->R_eq * r * p->Omega * sin(th) / (C_SI * sqrt(1.0 - rsr));
        w = 1.0 / sqrt(1.0 - beta * beta);
    }
    return w * dphi * r * r * sqrt(1.0 + f * f) * sin(th);
}

/* mesh_tools.pyx:391-426 */
static double theta_given_phi(double ph, double colat, double rad,
                              const double iv[2])
{
    double x = sin(colat) * sin(ph) / sin(rad);
    if (!(x <= 1.0)) return -1.0;
    x = asin(x);
    double th = tan(0.5 * (colat - rad)) / sin(0.5 * (x - ph));
    th *= sin(0.5 * (x + ph));
    th = atan(th);
    th = (th < 0.0) ? -2.0 * th : 2.0 * th;
    if (!(iv[0] <= th && th <= iv[1])) {
        x = M_PI - x;
        th = tan(0.5 * (colat - rad)) / sin(0.5 * (x - ph));
        th *= sin(0.5 * (x + ph));
        th = atan(th);
        th = (th < 0.0) ? -2.0 * th : 2.0 * th;
        if (!(iv[0] <= th && th <= iv[1])) return -1.0;
    }
    return th;
}

/* boundary-crossing azimuth of a circle on a parallel, folded into the cell
   (mesh_tools.pyx:524-620 top/bottom logic) */
static double cede_edge_phi(double th, double colat, double cedeR,
                            double pa, double pb)
{
    double v = eval_phi(th, colat, cedeR);
    if (v >= 0.0) {
        if (!(pa <= v && v <= pb)) {
            if (!(pa <= -v && -v <= pb)) v = -1.0;
            else if (v < 0.0) v = -v;
        } else if (v < 0.0) v = -v;
    }
    return v;
}

static double super_edge_phi(double th, double sColat, double sR, double sAzi,
                             double pa, double pb)
{
    double v;
    if (fabs(th - sColat) > sR) return -1.0;
    v = eval_phi(th, sColat, sR);
    if (v >= 0.0) {
        v += sAzi;
        if (v > M_PI) v -= TWO_PI;
        if (v < -M_PI) v += TWO_PI;
        if (!(pa <= v && v <= pb)) v = -1.0;
        else if (v < 0.0) v = -v;
    }
    return v;
}

static double super_edge_phi_neg(double th, double sColat, double sR,
                                 double sAzi, double pa, double pb,
                                 int strict)
{
    double v = eval_phi(th, sColat, sR);
    int take = strict ? (v > 0.0) : (v >= 0.0);
    if (take) {
        v = -v + sAzi;
        if (v > M_PI) v -= TWO_PI;
        if (v < -M_PI) v += TWO_PI;
        if (!(pa <= v && v <= pb)) v = -1.0;
        else if (v < 0.0) v = -v;
    }
    return v;
}

static double pick_first(double *c0, double *c1, double *c2, double *c3,
                         double ta, double tb)
{
    /* the first available crossing, consumed once it is used */
    if (*c0 >= 0.0) { double v = *c0; *c0 = -1.0; return v; }
    if (*c1 >= 0.0) { double v = *c1; *c1 = -1.0; return v; }
    if (*c2 >= 0.0) { *c2 = -1.0; return ta; }
    if (*c3 >= 0.0) { *c3 = -1.0; return tb; }
    return -1.0;
}

/* colatitudes t where the circle of angular radius rad about (TH, azimuth 0)
   meets the meridian at azimuth ph: A cos t + B sin t = C */
static void circle_meridian(double TH, double rad, double ph, double *t)
{
    double A = cos(TH), B = sin(TH) * cos(ph), C = cos(rad);
    double Rn = sqrt(A * A + B * B);
    t[0] = t[1] = -1.0;
    if (Rn == 0.0 || fabs(C) > Rn) return;
    double d = atan2(B, A), w = acos(C / Rn);
    t[0] = d - w;
    t[1] = d + w;
}

/*
 * The cell integrand has kinks where a region boundary crosses a cell edge
 * and square-root ends at a circle's extreme colatitudes.  The integral is
 * taken as the sum over the pieces between those points (an exact
 * decomposition: the rule then converges on each piece instead of bisecting
 * toward the kinks, so the result is accurate well below the tolerance).
 */
static int cell_breakpoints(double ta, double tb, double pa, double pb,
                            double colat, double cedeR, double superR,
                            double superAzi, double superColat, double *bp)
{
    int n = 0;
    double t2[2], cand[14];
    int nc = 0;
    circle_meridian(colat, cedeR, pa, t2); cand[nc++] = t2[0]; cand[nc++] = t2[1];
    circle_meridian(colat, cedeR, pb, t2); cand[nc++] = t2[0]; cand[nc++] = t2[1];
    cand[nc++] = colat - cedeR; cand[nc++] = colat + cedeR; cand[nc++] = cedeR - colat;
    if (superR > 0.0) {
        circle_meridian(superColat, superR, pa - superAzi, t2); cand[nc++] = t2[0]; cand[nc++] = t2[1];
        circle_meridian(superColat, superR, pb - superAzi, t2); cand[nc++] = t2[0]; cand[nc++] = t2[1];
        cand[nc++] = superColat - superR; cand[nc++] = superColat + superR;
        cand[nc++] = superR - superColat;
    }
    for (int k = 0; k < nc; k++)
        if (cand[k] > ta && cand[k] < tb) bp[n++] = cand[k];
    for (int q = 1; q < n; q++) {
        double v = bp[q];
        int r = q - 1;
        while (r >= 0 && bp[r] > v) { bp[r + 1] = bp[r]; r--; }
        bp[r + 1] = v;
    }
    return n;
}

static double quad_pieces(region_par *p, double lo, double hi, const double *bp, int nbp)
{
    double total = 0.0, a = lo;
    for (int k = 0; k <= nbp; k++) {
        double b = hi;
        if (k < nbp) {
            if (!(bp[k] > a && bp[k] < hi)) continue;
            b = bp[k];
        }
        total += mesh_quad(region_integrand, p, a, b);
        a = b;
    }
    return total;
}

/* mesh_tools.pyx:429-725 */
static double integrate_cell(double ta, double tb, double pa, double pb,
                             double R_eq, double eps, double zeta, int shape,
                             double colat, double cedeR, double superR,
                             double superAzi, double superColat)
{
    region_par p = {eps, zeta, colat, cedeR, superR, superAzi, superColat,
                    pa, pb, R_eq, 0.0, shape, 0, 1};
    double bp[14];
    int nbp = cell_breakpoints(ta, tb, pa, pb, colat, cedeR, superR, superAzi, superColat, bp);
    double integral = quad_pieces(&p, ta, tb, bp, nbp);
    if (eq12(integral, 0.0)) {
        double iv[2] = {ta, tb};
        double tl0 = theta_given_phi(pa, colat, cedeR, iv);
        double tl1 = theta_given_phi(pa - superAzi, superColat, superR, iv);
        double tr0 = theta_given_phi(pb, colat, cedeR, iv);
        double tr1 = theta_given_phi(pb - superAzi, superColat, superR, iv);
        double top0 = cede_edge_phi(ta, colat, cedeR, pa, pb);
        double top1 = super_edge_phi(ta, superColat, superR, superAzi, pa, pb);
        if (!(fabs(ta - superColat) > superR) && top1 < 0.0)
            top1 = super_edge_phi_neg(ta, superColat, superR, superAzi, pa, pb, 1);
        double bot0 = cede_edge_phi(tb, colat, cedeR, pa, pb);
        double bot1 = super_edge_phi(tb, superColat, superR, superAzi, pa, pb);
        if (!(fabs(tb - superColat) > superR) && bot1 < 0.0)
            bot1 = super_edge_phi_neg(tb, superColat, superR, superAzi, pa, pb, 0);
        double th[4];
        th[0] = pick_first(&tl0, &tr0, &top0, &bot0, ta, tb);
        th[1] = pick_first(&tl0, &tr0, &top0, &bot0, ta, tb);
        th[2] = pick_first(&tl1, &tr1, &top1, &bot1, ta, tb);
        th[3] = pick_first(&tl1, &tr1, &top1, &bot1, ta, tb);
        double tmin = tb, tmax = ta;
        for (int i = 0; i < 4; i++) {
            if (th[i] >= 0.0) {
                if (th[i] < tmin) tmin = th[i];
                if (th[i] > tmax) tmax = th[i];
            }
        }
        if (tmin < tmax) integral = quad_pieces(&p, tmin, tmax, bp, nbp);
    }
    return R_eq * R_eq * integral;
}

/* mesh_tools.pyx:816-862 */
static double integrate_spot(double ta, double tb, double R_eq, double eps,
                             double zeta, int shape, double colat,
                             double cedeR, double superR, double superColat,
                             double superAzi, int lorentz, double Omega)
{
    region_par p = {eps, zeta, colat, cedeR, superR, superAzi, superColat,
                    0.0, 0.0, R_eq, Omega, shape, lorentz, 0};
    return R_eq * R_eq * mesh_quad(region_integrand, &p, ta, tb);
}

static void clamp_even(double *s, size_t lo, size_t hi)
{
    if (*s < (double)lo) *s = (double)lo;
    else if (*s > (double)hi) *s = (double)hi;
    if (((size_t)(*s)) % 2 != 0) *s += 1.0;
}

/* mesh_tools.pyx:892-1099 */
void or_allocate_cells(size_t num_cells, size_t min_sqrt, size_t max_sqrt,
                       double mass, double r_s, double R_eq, double Omega,
                       double zeta, double eps, int shape, double superR,
                       double cedeR, double superColat, double cedeColat,
                       double superAzi, double holeColat, double holeR,
                       double holeAzi, double *out)
{
    (void)mass; (void)r_s;
    double sColat0 = superColat, s_lims[2], s_cell, s_area, s_sqrt;
    if (superColat - superR < 0.0 || superColat + superR > M_PI) {
        s_lims[0] = 0.0;
        if (superColat + superR > M_PI) {
            s_lims[1] = M_PI - superColat + superR;
            superColat = M_PI - superColat;
            holeColat = M_PI - holeColat;
        } else s_lims[1] = superColat + superR;
        s_cell = TWO_PI;
    } else {
        s_lims[0] = superColat - superR;
        s_lims[1] = superColat + superR;
        s_cell = 2.0 * asin(sin(superR) / sin(superColat));
    }
    s_cell *= integrate_area(s_lims[0], s_lims[1], R_eq, eps, zeta, shape, 0);
    s_area = integrate_spot(s_lims[0], s_lims[1], R_eq, eps, zeta, shape,
                            superColat, superR, holeR, holeColat, holeAzi, 0,
                            Omega);
    if (eq12(cedeR, 0.0)) {
        if (eq12(s_area, 0.0)) s_area = s_cell / 1000.0;
        s_sqrt = ceil(sqrt((double)num_cells * s_cell / s_area));
        clamp_even(&s_sqrt, min_sqrt, max_sqrt);
        out[0] = s_sqrt; out[1] = s_sqrt * s_sqrt; out[2] = 0.0; out[3] = 0.0;
        return;
    }
    double c_lims[2], c_cell, c_area, c_sqrt;
    if (cedeColat - cedeR < 0.0 || cedeColat + cedeR > M_PI) {
        c_lims[0] = 0.0;
        if (cedeColat + cedeR > M_PI) {
            c_lims[1] = M_PI - cedeColat + cedeR;
            cedeColat = M_PI - cedeColat;
            superColat = M_PI - sColat0;
        } else {
            c_lims[1] = cedeColat + cedeR;
            superColat = sColat0;
        }
        c_cell = TWO_PI;
    } else {
        c_lims[0] = cedeColat - cedeR;
        c_lims[1] = cedeColat + cedeR;
        superColat = sColat0;
        c_cell = 2.0 * asin(sin(cedeR) / sin(cedeColat));
    }
    c_cell *= integrate_area(c_lims[0], c_lims[1], R_eq, eps, zeta, shape, 0);
    c_area = integrate_spot(c_lims[0], c_lims[1], R_eq, eps, zeta, shape,
                            cedeColat, cedeR, superR, superColat, superAzi, 0,
                            Omega);
    if (eq12(s_area, 0.0)) s_area = s_cell / 1000.0;
    if (eq12(c_area, 0.0)) c_area = c_cell / 1000.0;
    double y = c_area / s_area - 1.0, s_num, c_num;
    if (eq12(y, 0.0)) {
        s_num = 0.5 * (double)num_cells;
        c_num = 0.5 * (double)num_cells;
    } else {
        s_num = (double)num_cells * ((sqrt(1.0 + y) - 1.0) / y);
        c_num = (double)num_cells - s_num;
    }
    s_sqrt = ceil(sqrt(s_num * s_cell / s_area));
    c_sqrt = ceil(sqrt(c_num * c_cell / c_area));
    clamp_even(&s_sqrt, min_sqrt, max_sqrt);
    clamp_even(&c_sqrt, min_sqrt, max_sqrt);
    out[0] = s_sqrt; out[1] = s_sqrt * s_sqrt;
    out[2] = c_sqrt; out[3] = c_sqrt * c_sqrt;
}

/* numpy.linspace element k: k*step + start, last element exactly stop */
static double lin(double start, double stop, size_t num, size_t k)
{
    if (num == 1) return start;
    if (k == num - 1) return stop;
    double step = (stop - start) / (double)(num - 1);
    return (double)k * step + start;
}

/*
 * mesh.pyx:23-417 (polar == 0) and polar_mesh.pyx:23-430 (polar == 1).
 * "cede" names the region meshed, "super" the region cut out of it.
 */
int or_construct_cell_mesh(int polar, size_t numCell, size_t S, double mass,
                           double r_s, double R_eq, double zeta, double eps,
                           int shape, double cedeR, double cedeColat,
                           double superR, double superColat, double superAzi,
                           double *theta, double *phi, double *rr,
                           double *cellAreas, double *maxAlpha,
                           double *cos_gamma, double *effGrav)
{
    (void)mass; (void)r_s;
    enum { NTAB = 1000 };
    double lims[2], boundary_phi = 0.0, cellArea, delta_phi;
    int invert = 0;
    if (polar) {
        lims[0] = 0.0;
        if (cedeColat + cedeR > M_PI) {
            invert = 1;
            lims[1] = M_PI - cedeColat + cedeR;
            cedeColat = M_PI - cedeColat;
            superColat = M_PI - superColat;
        } else lims[1] = cedeColat + cedeR;
        cellArea = TWO_PI * integrate_area(0.0, lims[1], R_eq, eps, zeta, shape, 0);
        delta_phi = TWO_PI / (double)S;
    } else {
        lims[0] = cedeColat - cedeR;
        lims[1] = cedeColat + cedeR;
        boundary_phi = asin(sin(cedeR) / sin(cedeColat));
        cellArea = 2.0 * boundary_phi;
        cellArea *= integrate_area(lims[0], lims[1], R_eq, eps, zeta, shape, 0);
        delta_phi = 2.0 * boundary_phi / (double)S;
    }
    double *phis = (double *)malloc(S * sizeof(double));
    if (polar) {
        for (size_t j = 0; j < S; j++)
            phis[j] = lin(-M_PI + 0.5 * delta_phi, M_PI - 0.5 * delta_phi, S, j);
    } else {
        for (size_t j = 0; j < S; j++)
            phis[j] = lin(-boundary_phi + 0.5 * delta_phi,
                          boundary_phi - 0.5 * delta_phi, S, j);
    }
    double *colTab = (double *)malloc(NTAB * sizeof(double));
    double *areaTab = (double *)malloc(NTAB * sizeof(double));
    for (size_t i = 0; i < NTAB; i++) colTab[i] = lin(lims[1], lims[0], NTAB, i);
    cellArea /= (double)numCell;
    double eta = cellArea / delta_phi;
    for (size_t i = 0; i < NTAB; i++)
        areaTab[i] = integrate_area(colTab[i], lims[1], R_eq, eps, zeta, shape, 0);
    double *sa = (double *)malloc(4 * NTAB * sizeof(double));
    or_steffen_init(areaTab, colTab, NTAB, sa, sa + NTAB, sa + 2 * NTAB, sa + 3 * NTAB);
    double *par = (double *)malloc((S > 1 ? S - 1 : 1) * sizeof(double));
    double i_eta = (double)S * eta;
    for (size_t i = 0; i + 1 < S; i++) {
        i_eta -= eta;
        par[i] = or_steffen_eval(areaTab, NTAB, sa, sa + NTAB, sa + 2 * NTAB,
                                 sa + 3 * NTAB, i_eta);
    }
    double s1[2] = {0, 0}, s2[2] = {0, 0}, s3[2] = {0, 0}, s4[2] = {0, 0};
    if (superR > 0.0) {
        if (superColat - superR < 0.0) {
            s1[0] = superR - superColat;
            s1[1] = superAzi + M_PI;
            if (s1[1] > M_PI) s1[1] -= TWO_PI;
            s2[0] = superColat + superR;
            s2[1] = superAzi;
        } else if (superColat + superR > M_PI) {
            s1[0] = superColat - superR;
            s1[1] = superAzi;
            s2[0] = TWO_PI - superColat - superR;
            s2[1] = superAzi + M_PI;
            if (s2[1] > M_PI) s2[1] -= TWO_PI;
        } else {
            s1[0] = superColat - superR;
            s1[1] = superAzi;
            s2[0] = acos(cos(superColat) / cos(superR));
            s2[1] = superAzi + asin(sin(superR) / sin(superColat));
            if (s2[1] > M_PI) s2[1] -= TWO_PI;
            s3[0] = superColat + superR;
            s3[1] = superAzi;
            s4[0] = s2[0];
            s4[1] = superAzi - asin(sin(superR) / sin(superColat));
            if (s4[1] < -M_PI) s4[1] += TWO_PI;
        }
    }
    int mirror = eq12(superColat, cedeColat) && eq12(superAzi, 0.0);
    size_t j_max = mirror ? S / 2 : S;
    memset(cellAreas, 0, S * S * sizeof(double));
    double *colat = (double *)malloc(S * sizeof(double));
    for (size_t i = 0; i < S; i++) {
        double lower, upper;
        if (i == S - 1) { lower = par[i - 1]; upper = lims[1]; }
        else if (i == 0) { lower = lims[0]; upper = par[i]; }
        else { lower = par[i - 1]; upper = par[i]; }
        colat[i] = integrate_area(lower, upper, R_eq, eps, zeta, shape, 1) / eta;
        double mu = cos(colat[i]);
        double rn = radius_norm(mu, eps, zeta, shape);
        rr[i] = rn * R_eq;
        effGrav[i] = or_effective_gravity(mu, R_eq, zeta, eps, shape);
        double f = f_theta(mu, rn, eps, zeta, shape);
        cos_gamma[i] = 1.0 / sqrt(1.0 + f * f);
        maxAlpha[i] = 0.5 * M_PI + acos(cos_gamma[i]);
        int p1 = 0, p3 = 0;
        if (polar && i == 0) {
            p1 = p3 = (cedeColat <= cedeR);
            if (p1 == 1) { p1 = p1 * (superR <= superColat); p3 = p1; }
            else p1 = p3 = 2;
        }
        double left = polar ? -M_PI : -boundary_phi;
        for (size_t j = 0; j < j_max; j++) {
            double right = left + delta_phi;
            double mid = left + 0.5 * delta_phi;
            int p2, p4, p5;
            if (!polar || i > 0) {
                p1 = (eval_psi(lower, left, cedeColat) <= cedeR);
                p3 = (eval_psi(lower, right, cedeColat) <= cedeR);
            }
            p2 = (eval_psi(upper, left, cedeColat) <= cedeR);
            p4 = (eval_psi(upper, right, cedeColat) <= cedeR);
            p5 = (eval_psi(colat[i], mid, cedeColat) <= cedeR);
            if (!polar || i > 0) {
                p1 = p1 == 1 ? (superR <= eval_psi(lower, left - superAzi, superColat)) : 2;
                p3 = p3 == 1 ? (superR <= eval_psi(lower, right - superAzi, superColat)) : 2;
            }
            p2 = p2 == 1 ? (superR <= eval_psi(upper, left - superAzi, superColat)) : 2;
            p4 = p4 == 1 ? (superR <= eval_psi(upper, right - superAzi, superColat)) : 2;
            p5 = p5 == 1 ? (superR <= eval_psi(colat[i], mid - superAzi, superColat)) : 2;
            if (!(p1 == p2 && p2 == p3 && p3 == p4 && p4 == p5)) {
                cellAreas[i * S + j] = integrate_cell(lower, upper, left, right,
                        R_eq, eps, zeta, shape, cedeColat, cedeR, superR,
                        superAzi, superColat);
            } else if (superR > 0.0) {
#define INSIDE(s) ((lower <= (s)[0] && (s)[0] <= upper) && (left <= (s)[1] && (s)[1] <= right))
                if (INSIDE(s1) || INSIDE(s2) || INSIDE(s3) || INSIDE(s4)) {
                    cellAreas[i * S + j] = integrate_cell(lower, upper, left,
                            right, R_eq, eps, zeta, shape, cedeColat, cedeR,
                            superR, superAzi, superColat);
                } else if (p5 == 1) cellAreas[i * S + j] = cellArea;
#undef INSIDE
            } else if (p5 == 1) cellAreas[i * S + j] = cellArea;
            if (mirror) cellAreas[i * S + (S - 1 - j)] = cellAreas[i * S + j];
            left = right;
        }
    }
    for (size_t i = 0; i < S; i++)
        for (size_t j = 0; j < S; j++) {
            phi[i * S + j] = phis[j];
            theta[i * S + j] = invert ? M_PI - colat[i] : colat[i];
        }
    free(phis); free(colTab); free(areaTab); free(sa); free(par); free(colat);
    return 0;
}

/*
 * global_mesh.pyx:21-147 (construct_closed_cellMesh): equal-area parallels
 * over the whole star, S rings of S cells, every cell of area *cellArea.
 * The northern S/2 ring colatitudes are the area-weighted mean colatitudes
 * between parallels found by inverting a 1000-point area table (Steffen);
 * the southern rings mirror them.
 */
int or_construct_closed_mesh(size_t S, double R_eq, double zeta, double eps,
                             int shape, double *colat, double *phis,
                             double *rr, double *cellArea, double *maxAlpha,
                             double *cos_gamma, double *effGrav)
{
    enum { NTAB = 1000 };
    const double hpi = 0.5 * M_PI;
    size_t ncc = S / 2;
    if (ncc < 2) return 1;
    double ca = TWO_PI * integrate_area(0.0, hpi, R_eq, eps, zeta, shape, 0) / (double)(S * S);
    ca *= 2.0;
    *cellArea = ca;
    double eta = ca * (double)S / TWO_PI;
    double *colTab = (double *)malloc(NTAB * sizeof(double));
    double *areaTab = (double *)malloc(NTAB * sizeof(double));
    double *sa = (double *)malloc(4 * NTAB * sizeof(double));
    double *par = (double *)malloc(ncc * sizeof(double));
    for (size_t i = 0; i < NTAB; i++) colTab[i] = lin(hpi, 0.0, NTAB, i);
    for (size_t i = 0; i < NTAB; i++)
        areaTab[i] = integrate_area(colTab[i], hpi, R_eq, eps, zeta, shape, 0);
    or_steffen_init(areaTab, colTab, NTAB, sa, sa + NTAB, sa + 2 * NTAB, sa + 3 * NTAB);
    double i_eta = (double)ncc * eta;
    for (size_t i = 0; i + 1 < ncc; i++) {
        i_eta -= eta;
        par[i] = or_steffen_eval(areaTab, NTAB, sa, sa + NTAB, sa + 2 * NTAB,
                                 sa + 3 * NTAB, i_eta);
    }
    for (size_t i = 0; i < ncc; i++) {
        double lo, hi;
        if (i == ncc - 1) { lo = par[i - 1]; hi = hpi; }
        else if (i == 0) { lo = 0.0; hi = par[0]; }
        else { lo = par[i - 1]; hi = par[i]; }
        colat[i] = integrate_area(lo, hi, R_eq, eps, zeta, shape, 1) / eta;
        double mu = cos(colat[i]);
        double rn = radius_norm(mu, eps, zeta, shape);
        rr[i] = rn * R_eq;
        effGrav[i] = or_effective_gravity(mu, R_eq, zeta, eps, shape);
        double f = f_theta(mu, rn, eps, zeta, shape);
        cos_gamma[i] = 1.0 / sqrt(1.0 + f * f);
        maxAlpha[i] = hpi + acos(cos_gamma[i]);
    }
    double dphi = TWO_PI / (double)S;
    for (size_t j = 0; j < S; j++) phis[j] = lin(-M_PI + 0.5 * dphi, M_PI - 0.5 * dphi, S, j);
    for (size_t i = ncc; i < S; i++) {
        colat[i] = M_PI - colat[S - i - 1];
        rr[i] = rr[S - i - 1];
        cos_gamma[i] = cos_gamma[S - i - 1];
        effGrav[i] = effGrav[S - i - 1];
        maxAlpha[i] = maxAlpha[S - i - 1];
    }
    free(colTab); free(areaTab); free(sa); free(par);
    return 0;
}
