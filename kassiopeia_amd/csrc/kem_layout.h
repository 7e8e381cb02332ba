/*
 * kem_layout.h -- device-resident element record layout (shared by host preparation and
 * HIP kernels).
 *
 * Each boundary element is expanded once, on the host, from KEMField's 11-double shape
 * record (KElectrostaticCubatureTriangleIntegrator.cc:457-467: [A, B, P0, N1, N2]; line
 * segments [P0, P1, diameter], KElectrostaticQuadratureLineSegmentIntegrator.cc:197-203)
 * into a fixed-stride FP64 record holding every per-element quantity the integrals need:
 * dispatch centroid and 1/size, cubature prefactor, plane normal, vertices, and the
 * per-edge quantities of the closed-form (solid angle + edge logarithm) near field.  The
 * kernels read these records with wave-uniform loads; nothing per element is recomputed
 * per (point, element) pair except what depends on the point.
 *
 * Cubature nodes are held in a second array, KEM_NODE_SLOTS nodes of 4 doubles (x, y, z, w)
 * per element: triangle/rectangle 7-point at slot 0, 12-point at slot 7, 33-point at slot 19;
 * line segments 4-node Gauss-Legendre at slot 0 and 16-node at slot 4 (positions of both
 * members of each symmetric pair, weight * half-length).
 */
#ifndef KEM_LAYOUT_H
#define KEM_LAYOUT_H

/* conic sections (surface of revolution of P0-P1 about z) and rings (a charged circle) are the
   reference's axisymmetric shapes, KConicSection.hh / KRing.hh */
enum kem_shape_type { KEM_TRIANGLE = 0, KEM_RECTANGLE = 1, KEM_LINESEGMENT = 2, KEM_CONICSECTION = 3, KEM_RING = 4 };

enum kem_geom_field {
    KG_CX = 0, /* dispatch centre (triangle centroid, rectangle centre, segment midpoint) */
    KG_CY = 1,
    KG_CZ = 2,
    KG_INVSIZE = 3, /* 1 / size: mean side (triangle), (A+B)/2 (rectangle), length (segment) */
    KG_PREF = 4,    /* area/(4 pi eps0) (triangle, rectangle); d/(4 eps0) (segment) */
    KG_TYPE = 5,    /* shape type as a double */
    KG_NX = 6,      /* unit plane normal */
    KG_NY = 7,
    KG_NZ = 8,
    KG_HX = 9, /* centroid used for the height h = n.(P - c) of the closed form */
    KG_HY = 10,
    KG_HZ = 11,
    KG_NE = 12, /* number of edges (3, 4) */
    KG_A = 13,  /* side A (rectangle in-plane test); segment, conic section: length L */
    KG_B = 14,  /* side B; segment: diameter */
    KG_V = 16,  /* vertices V0..V3 (segment, conic section: P0 at V0, P1 at V1; ring: P at V0) */
    KG_U = 28,  /* unit edge directions u_e (segment: unit axis at u_0) */
    KG_MID = 40, /* edge midpoints */
    KG_M = 52,   /* in-plane outward edge normals m_e = u_e x n */
    KG_SV = 64,  /* full edge vectors (point-on-edge detection) */
    KG_IL = 76,  /* 1 / edge length */
    /* squared distance thresholds |P - centre|^2 = (dr * size)^2 of the rule ladders:
       near set (closed form), reference ladder (7- vs 12-point, 12- vs 33-point; segments
       4- vs 16-node), mixed-precision FP32 ladder (7- vs 12-point, 12- vs 33-point) */
    KG_NEAR2 = 80,
    KG_REF7_2 = 81,
    KG_REF12_2 = 82,
    KG_MIX7_2 = 83,
    KG_MIX12_2 = 84,
    KG_MIX6_2 = 85, /* mixed ladder: triangle 6-point rule beyond this (no other shape) */
    KG_STRIDE = 88
};

/* raw shape record kept for the analytic / reference policies (11 doubles, padded) */
#define KEM_REC 12

#define KEM_NODE_SLOTS 52
#define KEM_NODE_STRIDE (4 * KEM_NODE_SLOTS)
/* the assembly's potential node table (Dirichlet rows, numeric policy; built on the device per
   assembly): per element KEM_NODE_SLOTS nodes of 5 doubles relative to the element centre c,
   (-2 q'x, -2 q'y, -2 q'z, |q'|^2, w) with q' = Q - c, so that |P - Q|^2 = |P - c|^2 + |q'|^2 -
   2 (P - c) . q' costs 4 operations per node instead of 6 */
#define KEM_ANODE_STRIDE (5 * KEM_NODE_SLOTS)
#define KEM_N7 0
#define KEM_N12 7
#define KEM_N33 19
#define KEM_L4 0
#define KEM_L16 4
/* the FP32 node table (mixed far field) holds 6 more slots: a degree-4 6-point triangle rule */
#define KEM_N6 52
#define KEM_NODE32_SLOTS 58
#define KEM_NODE32_STRIDE (4 * KEM_NODE32_SLOTS)

/* element clusters for the near pass: elements in Morton order of their centres, KEM_CLUSTER
   per cluster; each cluster keeps (centre x, y, z, radius) with radius = max over its elements
   of |c_j - centre| + near distance of j, so a point farther than radius from the centre has no
   near pair in the cluster */
#define KEM_CLUSTER 32

/* far records of the mixed far field, one per element in cluster order (32 floats = 128 B):
   the 6-point rule's nodes as (x, y, z, w) FP32 offsets from the CLUSTER centre (the element's
   node offsets, slots KEM_N6.. of the FP32 node table, plus element centre - cluster centre,
   rounded once), then the element centre and the 6-point reach sqrt((12 size)^2) * (1 + 1e-9)
   as FP64 (infinite for other shapes).  A point farther than the reach from the element
   centre is a 6-point pair.  Per cluster: centre x, y, z and the largest |c_j - centre| +
   reach_j (FP64), so a point farther than that from the centre is a 6-point pair with every
   element of the cluster. */
#define KEM_FAR6_STRIDE 32
#define KEM_FAR6_CENTRE 24
#define KEM_FAR6_REACH 30

/* 3-point far records, one per element in cluster order (16 floats = 64 B): the degree-2 rule
   with nodes at barycentric (2/3, 1/6, 1/6) and its permutations, weights 1/3, as (x, y, z, 0)
   FP32 offsets from the cluster centre (computed in FP64, rounded once), for triangles beyond
   KEM_FAR3_RATIO sizes, where its truncation (<= 3.0e-8 relative, measured) is on the level of
   the mixed ladder's other rules.  Zero records for other shapes. */
#define KEM_FAR3_STRIDE 16
#define KEM_FAR3_RATIO 80.0

/* per-cluster far data (doubles): centre x, y, z, 6-point reach, 3-point reach (the largest
   |c_j - centre| + KEM_FAR3_RATIO size_j over the cluster, infinite when it holds any other
   shape than triangles), 3 pad */
#define KEM_FAR_CL_STRIDE 8

/* collocation target record (one per element): centre, unit normal, boundary type, flux */
enum kem_target_field { KT_CX = 0, KT_CY, KT_CZ, KT_NX, KT_NY, KT_NZ, KT_BC, KT_FLUX, KT_STRIDE = 8 };

#endif
