// Plain-old-data descriptors shared by host launch code and kernels.
#pragma once
#include <stdint.h>

namespace xh {

// One meshed member of a hot region: the superseding region (member 0) or
// the ceding ring (member 1).  "cede*" names the region that is meshed and
// "super*" the region cut out of it (the reference's mesh convention).
struct Member {
    // spacetime
    double R_eq, epsilon, zeta, Omega, r_s, inclination, mode_freq, d_sq;
    int shape;          // 0 = oblate (AlGendy & Morsink 2014), 1 = sphere
    int active;         // member present (ceding ring only when its radius > 0)
    // mesh geometry
    double cedeR, cedeColat, superR, superColat, superAzi;
    double phi_offset;  // added to cell azimuths after meshing
    int polar;          // mesh centred on a pole
    int S;              // cells per side, written by the allocation kernel
    double numCell;
    // radiation field
    double logT;        // log10 effective temperature of the member
    int status;         // 0 ok, 1 ray failure, 2 pulse failure, 3/4 size limits
    int antiphased;     // add pi to azimuths after the offset
    // elsewhere correction: subtract the blackbody intensity at else_logT
    // from every cell and keep negative spline values (correct != 0)
    double else_logT;
    int correct;
};

// Inputs of allocate_cells for one hot region (mesh_tools.pyx:892-910).
struct RegionAlloc {
    double superR, cedeR, superColat, cedeColat, superAzi;
    double holeColat, holeR, holeAzi;
    double R_eq, Omega, zeta, epsilon;
    int shape, num_cells, min_sqrt, max_sqrt;
    int member0, member1;   // indices into the Member array (member1 < 0: none)
};

// Fixed per-member buffer strides (elements).
struct Layout {
    int S_max;      // max cells per side
    int N_R;        // rays per parallel
    int N_L;        // leaves
    int N_P;        // model phases
    int N_E;        // energies
    int IO_max;     // max image orders traced
    int image_limit;// 0 = infer from max deflection
    int atm_ext;    // 1 blackbody, 2 four-dimensional table
};

}  // namespace xh
