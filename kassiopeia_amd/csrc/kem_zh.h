// kem_zh.h -- zonal-harmonic expansion of axisymmetric elements: host-side tree and
// coefficients (kem_zh_gen.cpp), device evaluation (kem_zh.hip).
#ifndef KEM_ZH_H
#define KEM_ZH_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace kem {

// KZonalHarmonicParameters defaults (FieldSolvers/ZonalHarmonic/Generator/include/
// KZonalHarmonicParameters.hh:11-27)
struct ZHParams {
    int nbif = 5;
    double conv_ratio = 0.99;
    double prox = 1.e-14;
    double conv_param = 1.e-15;
    int ncen = 500;
    bool frac_spacing = true;
    double frac_dist = 0.2;
    double cen_dz = 0.01;
    double cen_z1 = 0., cen_z2 = 0.;
    int nrem = 500;
    int nrem_sp = 3;
    double rem_z1 = 0., rem_z2 = 0.;
};

// one element as the generators see it: a ring (radius r0 at z0, total charge sigma), a conic
// section (r0, z0) -> (r1, z1) with surface charge density sigma, or an element that is not
// axisymmetric (integrated directly)
struct ZHElement {
    enum { kRing = 0, kConic = 1, kOther = 2 };
    int type = kOther;
    double r0 = 0., z0 = 0., r1 = 0., z1 = 0.;
    double sigma = 0.;
};

struct ZHSourcePoint {
    double z0 = 0., rho = 0.;
    int64_t offset = 0;  // first coefficient in ZHTables::coeff
    int ncoef = 0;
};

// the flattened tree: node 0 is the head; per node 8 ints [central begin, count, remote begin,
// count, children begin (in child), count, elements begin (in elem), count]
struct ZHTables {
    std::vector<int32_t> node, child;
    std::vector<int64_t> elem;
    std::vector<ZHSourcePoint> sp;
    std::vector<double> coeff;
    std::vector<double> leg;  // 4 rows of nleg recursion coefficients
    int nleg = 0;
    int depth = 1;
};

ZHTables zh_build(const std::vector<ZHElement>& elements, const ZHParams& params);

// device view of a ZHTables upload
struct ZHDevice {
    const int32_t* node = nullptr;
    const int32_t* child = nullptr;
    const int64_t* elem = nullptr;
    const double* sp = nullptr;  // per source point 4 doubles: z0, rho, coefficient offset, count
    const double* coeff = nullptr;
    const double* leg = nullptr;
    int nleg = 0;
    int depth = 1;
    double conv_ratio = 0.99, prox = 1.e-14, conv_param = 1.e-15;
};

// (E, phi) of the expansion at npts device points; the direct fallback integrates the node's
// elements (geom / nodes / recs of the mesh, sigma per element, policy of the context) with
// Kahan sums; counts (nullable, 3 zeroed counters): central, remote and direct evaluations
hipError_t launch_zh_field(hipStream_t st, const ZHDevice& z, const double* geom, const double* nodes,
                           const double* recs, const double* sigma, int policy, const double* px, const double* py,
                           const double* pz, int64_t npts, double* phi, double* ex, double* ey, double* ez,
                           unsigned long long* counts);

}  // namespace kem

#endif
