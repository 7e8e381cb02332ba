// kem_prep.h -- host-side element preparation and physical constants.
#ifndef KEM_PREP_H
#define KEM_PREP_H

#include "kem_layout.h"

#include <cstdint>

namespace kem {

// Kommon/Base/Utility/KConst_2021.h:37 (pi), :139-146 (eps0); KEMField/Source/Core/src/KEMConstants.cc:9-12
constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kEps0 = 8.8541878128E-12;
constexpr double kOneOverFourPiEps0 = 1. / (4. * kPi * kEps0);

// Expand one KEMField shape record into the device record g[KG_STRIDE] and the node table
// nodes[KEM_NODE_STRIDE].  Returns 0, or -1 for a degenerate / non-finite element.
int prep_element(int type, const double* shape_data, double* g, double* nodes);

// Barycentric coordinates and weights of the triangle rules in node-slot order (7-, 12-,
// 33-point at KEM_N7 / KEM_N12 / KEM_N33): t[4*slot .. 4*slot+3] = (a, b, c, w), the node
// being a*P0 + b*P1 + c*P2 (GaussPoints_Tri7P/12P/33P, KElectrostaticCubatureTriangleIntegrator.cc:34-270).
void triangle_rule_table(double* t);

// FP32 node table for the mixed-precision far field: node positions relative to the
// element's dispatch centre (KG_CX..KG_CZ of g) and weights, KEM_NODE32_STRIDE floats (the
// 52 slots of the FP64 table plus a 6-point triangle rule at KEM_N6).
void prep_nodes32(const double* g, const double* nodes, float* nodes32);

// Collocation target record t[KT_STRIDE]: the shape's Centroid(), unit normal, boundary type
// and Neumann flux.
void prep_target(int type, const double* shape_data, int bc, double flux, double* t);

// Element clusters for the near pass: order[n] lists the elements in Morton order of their
// dispatch centres; clusters[4 * ceil(n / KEM_CLUSTER)] holds, per run of KEM_CLUSTER
// consecutive entries of order, the centre of their bounding box and a radius R such that
// every point within the near distance of a member lies within R of the centre.
void build_clusters(int64_t n, const double* g, int32_t* order, double* clusters);

// Far records of the mixed far field in cluster order (layout: KEM_FAR6_*, KEM_FAR3_* and
// KEM_FAR_CL_STRIDE in kem_layout.h): far6[KEM_FAR6_STRIDE * n] floats from the FP32 node table
// nodes32 and the element records g, far3[KEM_FAR3_STRIDE * n] floats from the vertices in g,
// far_cl[KEM_FAR_CL_STRIDE * ceil(n / KEM_CLUSTER)] doubles with the clusters' centres taken from
// clusters.
void build_far6(int64_t n, const double* g, const float* nodes32, const int32_t* order, const double* clusters,
                float* far6, float* far3, double* far_cl);

}  // namespace kem

#endif
