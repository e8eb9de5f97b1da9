// Launch interfaces of the fused MI355X kernels (host side of the backend calls these).
#pragma once

#include <hip/hip_runtime.h>

#include "honee_mi355x.h"

// Strided quadrature-data layout: value (elem e, component c, point i) at i*sn + c*sc + e*se.
struct QDataLayout {
  long long sn, sc, se;
};

// Device copies of the 1D tables of a tensor basis: interpolation B [Q][P] and collocated derivative Dc [Q][Q].
struct BasisTables {
  const double *B;
  const double *Dc;
  const double *BB, *GB, *GG;  // elementwise products B.B, G.B, G.G of the 1D tables (diagonal assembly)
  const double *G = nullptr;   // derivative of the interpolation [Q][P] (null: formed as Dc B)
};

struct SetupArgs {
  int           nelem;
  const int    *xoffsets;  // [nelem][Px^3], L index of component 0
  int           xcomp_stride;
  const double *x;
  const double *Bx, *Dx;  // [Q][Px]
  const double *w1d;      // [Q]
  double       *qdata;
  QDataLayout   qd;
};

struct VolumeArgs {
  int           nelem;
  const int    *offsets;  // [nelem][P^3], L index of component 0
  int           comp_stride;
  const double *u;        // active input L-vector (q or dq)
  const double *u2;       // second L-vector input (q_dot), if any
  double       *v;        // output L-vector (summed into with atomics when ev is null)
  double       *ev;       // two-pass scatter: element-node contributions [nelem * P^3][5] ([..][25] for point blocks; replaces v)
  float        *ev32;     // the same contributions in FP32 (element-ordered; the FP32 Jacobian apply), when ev is null
  const int    *perm;     // two-pass scatter: slot of element node e * P^3 + j in ev (null: slot e * P^3 + j)
  const double *qdata;
  QDataLayout   qd;
  double       *jac;      // FP64 jac_data written by the IFunction
  const double *jacr;     // FP64 jac_data read by the IJacobian
  float        *jac32;    // FP32 jac_data mirror (written by IFunction / read by IJacobian in FP32 mode)
  QDataLayout   jd;
  BasisTables   bt;
  const int    *xoffsets;  // coordinate restriction [nelem][8]
  const double *x;
  const double *Bx;        // [Q][2]
  // qdata recomputed from the coordinates (null: read qdata): element vertex coordinates [nelem][3][8] and the
  // coordinate basis tables {Bx [Q][2], Dx [Q][2], w [Q]}
  const double *xe;
  const double *gtab;
};

// Volume operator whose fields have their own restrictions (SGS closure inputs and residual, Mass_N): an active input
// u (in_offsets), a passive interpolated input u2 (in2_offsets), strided qdata, and the output v (out_offsets), or the
// two-pass E-vector ev with slots perm[] of the output restriction's transpose map.
struct FieldArgs {
  int           nelem;
  const int    *in_offsets;
  int           in_comp_stride;
  const double *u;
  const int    *in2_offsets;
  int           in2_comp_stride;
  const double *u2;
  const int    *out_offsets;
  int           out_comp_stride;
  double       *v;
  double       *ev;
  const int    *perm;
  const double *qdata;
  QDataLayout   qd;
  BasisTables   bt;
  const double *aux;  // per-point auxiliary data (collocated anisotropy tensor [elem][7][Q^3])
  QDataLayout   auxl;
  const int    *xoffsets;  // trilinear coordinate field: [nelem][8] offsets of component 0
  int           xcomp_stride;
  const double *x;
  const double *Bx;  // [Q][2]
};

// Nodal SGS closure: one evaluation per row k of the output restriction's transpose map (rowptr, index, row_offset),
// summed over the row's element nodes with their inverse multiplicities (strided layout im, elem_size nodes each).
struct SgsNodalArgs {
  long long     nrows;
  const int    *rowptr, *index, *row_offset;
  int           out_comp_stride;
  double       *v;
  const int    *q_offsets, *g_offsets, *a_offsets;
  int           q_comp_stride, g_comp_stride, a_comp_stride;
  const double *q, *g, *aniso;
  const double *inv_mult;
  QDataLayout   im;
  int           elem_size;
  const double *model;  // b1[20], b2[6], W1[20][6], W2[6][20], bounds[6][2], alpha
  const int    *p_offsets;  // training pairs: filtered velocity products (offsets into the filtered field q)
  int           p_comp_stride;
};

hipError_t launch_gather_vertex_coords(int nelem, const int *xoffsets, const double *x, int xcs, double *xe, hipStream_t s);
hipError_t launch_setup_qdata(int Px, int Q, const SetupArgs &a, hipStream_t stream);
bool       volume_pq_supported(int P, int Q);
hipError_t launch_ifunction(int P, int Q, int var, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t stream);
hipError_t launch_ijacobian(int P, int Q, int var, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t stream);
hipError_t launch_ijacobian_diagonal(int P, int Q, int var, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t stream);
// Point-block diagonals [node][c_out][c_in] (output index 25 * (offset / 5) + 5 c_out + c_in; interleaved L-vectors)
hipError_t launch_ijacobian_point_block_diagonal(int P, int Q, int var, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx,
                                                 hipStream_t stream);
hipError_t launch_mass_point_block_diagonal(int P, int Q, bool stabilized, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t stream);
hipError_t launch_rhs(int P, int Q, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t stream);
hipError_t launch_mass(int P, int Q, bool stabilized, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t stream);
hipError_t launch_mass_diagonal(int P, int Q, bool stabilized, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t stream);
hipError_t launch_monitor_ke(int P, int Q, int var, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t stream);

// Vector kernels (kernels_vector.hip)
// Boundary faces (2D tensor faces of the hexahedral mesh); VolumeArgs::nelem is the number of faces
bool       face_pq_supported(int P, int Q);
hipError_t launch_setup_boundary(int Q, const SetupArgs &a, hipStream_t stream);
hipError_t launch_face_boundary_integral(int P, int Q, int var, bool jacobian, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx,
                                         hipStream_t s);
hipError_t launch_face_freestream(int P, int Q, int var, bool hllc, bool jacobian, bool fp32, const VolumeArgs &a, const CeedMiFreestreamContext &ctx,
                                  hipStream_t s);
// Diagonal (point_block false) or point-block diagonal [node][c_out][c_in] of a face Jacobian
hipError_t launch_face_jacobian_diagonal(bool point_block, int P, int Q, int var, bool freestream, bool hllc, bool fp32, const VolumeArgs &a,
                                         const CeedMiNewtonianContext &nctx, const CeedMiFreestreamContext &fctx, hipStream_t s);
// Outflow (Riemann or weak pressure) and slip faces (kernels_face_bc.hip); kind of the diagonal: one of kFace*
constexpr int kFaceRiemannOutflow = 2, kFacePressureOutflow = 3, kFaceSlip = 4;
hipError_t launch_face_outflow(int P, int Q, int var, bool pressure, bool jacobian, bool fp32, const VolumeArgs &a, const CeedMiOutflowContext &ctx,
                               hipStream_t s);
hipError_t launch_face_slip(int P, int Q, int var, bool jacobian, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s);
hipError_t launch_face_bc_jacobian_diagonal(bool point_block, int P, int Q, int var, int kind, bool fp32, const VolumeArgs &a,
                                            const CeedMiNewtonianContext &nctx, const CeedMiOutflowContext &octx, hipStream_t s);

hipError_t launch_fill(double *x, double value, long long n, hipStream_t s);
hipError_t launch_axpby(double *y, double alpha, double beta, const double *x, long long n, hipStream_t s);
hipError_t launch_scale(double *x, double alpha, long long n, hipStream_t s);
hipError_t launch_pointwise_mult(double *w, const double *x, const double *y, long long n, hipStream_t s);
hipError_t launch_reciprocal(double *x, long long n, hipStream_t s);
hipError_t launch_norm(const double *x, long long n, int type, double *d_partial, int npartial, hipStream_t s);
hipError_t launch_multiplicity(const int *offsets, long long n_offsets, int ncomp, int comp_stride, double *mult, hipStream_t s);
hipError_t launch_restriction_notranspose(const int *offsets, int nelem, int elemsize, int ncomp, int comp_stride, const double *u, double *e, hipStream_t s);
// Second pass of the deterministic scatter: v[t_offset[k] + c * comp_stride] (+)= sum over j in
// [t_rowptr[k], t_rowptr[k+1]) of ev[slot(j) * ncomp + c] in ascending j, with slot(j) = t_index[j] for an
// element-ordered E-vector or slot(j) = j for a node-sorted one (t_index null).
hipError_t launch_node_gather(int nrows, int ncomp, int comp_stride, const int *t_rowptr, const int *t_index, const int *t_offset,
                              const double *ev, double *v, bool overwrite, hipStream_t s);
// The same for 5 x 5 point blocks: v[25 (t_offset[k] / 5) + i] += sum over the node's slots of ev[t_index[j] * 25 + i]
// Two-pass scatter with an FP32 element-ordered E-vector [elem node][5]: sums per node in FP64, in slot order
hipError_t launch_node_gather_rows_f32(int nrows, int comp_stride, const int *t_rowptr, const int *t_index, const int *t_offset, const float *ev,
                                      double *v, bool overwrite, hipStream_t s);
hipError_t launch_node_gather_blocks(int nrows, const int *t_rowptr, const int *t_index, const int *t_offset, const double *ev, double *v,
                                     hipStream_t s);
hipError_t launch_restriction_transpose(const int *offsets, int nelem, int elemsize, int ncomp, int comp_stride, const double *e, double *u, hipStream_t s);

// SGS closure and projections (kernels_sgs.hip)
bool       sgs_pq_supported(int P, int Q);
bool       mass_n_supported(int ncomp);
size_t     sgs_model_doubles();
hipError_t launch_velocity_gradient_rhs(int P, int Q, int var, const FieldArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s);
hipError_t launch_anisotropy_rhs(int P, int Q, const FieldArgs &a, hipStream_t s);
hipError_t launch_mass_n(int P, int Q, int ncomp, const FieldArgs &a, hipStream_t s);
hipError_t launch_mass_n_diagonal(int P, int Q, int ncomp, const FieldArgs &a, hipStream_t s);
hipError_t launch_ifunction_nodal_sgs(int P, int Q, int var, const FieldArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s);
hipError_t launch_inverse_multiplicity(const int *offsets, int nelem, int elem_size, const double *mult, double *out, QDataLayout ol, hipStream_t s);
hipError_t launch_sgs_nodal(int var, const SgsNodalArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s);
// Differential filter and SGS training pairs (kernels_sgs.hip)
hipError_t launch_diff_filter_rhs(int P, int Q, int var, const FieldArgs &a, const CeedMiDiffFilterContext &ctx, hipStream_t s);
bool       diff_filter_ncomp_supported(int ncomp);
hipError_t launch_diff_filter_lhs(int P, int Q, int ncomp, const FieldArgs &a, const CeedMiDiffFilterContext &ctx, hipStream_t s);
hipError_t launch_diff_filter_lhs_diagonal(int P, int Q, int ncomp, const FieldArgs &a, const CeedMiDiffFilterContext &ctx, hipStream_t s);
hipError_t launch_anisotropy_collocate(int Q, int nelem, const double *qdata, QDataLayout qd, double *out, QDataLayout ol, hipStream_t s);
hipError_t launch_sgs_training_nodal(const SgsNodalArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s);
// Deterministic dot product: npartial fixed per-block partial sums (summed in order by the caller)
hipError_t launch_dot(const double *x, const double *y, long long n, double *d_partial, int npartial, hipStream_t s);
