// Fused kernels of the outflow and slip boundary-face operators on MI355X: RiemannOutflow, PressureOutflow
// (qfunctions/bc_outflow.h:37-305) and Slip (qfunctions/bc_slip.h:11-111), each with its Jacobian (FP64 or FP32) and
// the Jacobi diagonal and point blocks of the Jacobian; instantiations of the face kernel templates in
// face_kernels.hpp.
#include "face_kernels.hpp"

namespace {

template <int P, int Q, class F, class JD, class JF, class Ctx>
hipError_t bc_apply(bool jacobian, bool fp32, const VolumeArgs &a, const Ctx &ctx, hipStream_t s) {
  if (!jacobian) return launch_face_op<P, Q, F>(a, ctx, s);
  return fp32 ? launch_face_op<P, Q, JF>(a, ctx, s) : launch_face_op<P, Q, JD>(a, ctx, s);
}

template <int P, int Q, template <int> class F, template <int, typename> class J, class Ctx>
hipError_t bc_apply_var(int var, bool jacobian, bool fp32, const VolumeArgs &a, const Ctx &ctx, hipStream_t s) {
  switch (var) {
    case 0: return bc_apply<P, Q, F<0>, J<0, double>, J<0, float>>(jacobian, fp32, a, ctx, s);
    case 1: return bc_apply<P, Q, F<1>, J<1, double>, J<1, float>>(jacobian, fp32, a, ctx, s);
    default: return bc_apply<P, Q, F<2>, J<2, double>, J<2, float>>(jacobian, fp32, a, ctx, s);
  }
}

template <int P, int Q, class JD, class JF, class Ctx>
hipError_t bc_diag(bool pblock, bool fp32, const VolumeArgs &a, const Ctx &ctx, hipStream_t s) {
  if (fp32) return pblock ? launch_face_pblock<P, Q, JF>(a, ctx, s) : launch_face_diag<P, Q, JF>(a, ctx, s);
  return pblock ? launch_face_pblock<P, Q, JD>(a, ctx, s) : launch_face_diag<P, Q, JD>(a, ctx, s);
}

template <int P, int Q, template <int, typename> class J, class Ctx>
hipError_t bc_diag_var(bool pblock, int var, bool fp32, const VolumeArgs &a, const Ctx &ctx, hipStream_t s) {
  switch (var) {
    case 0: return bc_diag<P, Q, J<0, double>, J<0, float>>(pblock, fp32, a, ctx, s);
    case 1: return bc_diag<P, Q, J<1, double>, J<1, float>>(pblock, fp32, a, ctx, s);
    default: return bc_diag<P, Q, J<2, double>, J<2, float>>(pblock, fp32, a, ctx, s);
  }
}

template <int P, int Q>
hipError_t outflow_pq(int var, bool pressure, bool jacobian, bool fp32, const VolumeArgs &a, const CeedMiOutflowContext &ctx, hipStream_t s) {
  if (pressure) return bc_apply_var<P, Q, fop::PressureOutflow, fop::PressureOutflowJacobian>(var, jacobian, fp32, a, ctx, s);
  return bc_apply_var<P, Q, fop::RiemannOutflow, fop::RiemannOutflowJacobian>(var, jacobian, fp32, a, ctx, s);
}

template <int P, int Q>
hipError_t bc_diag_pq(bool pblock, int var, int kind, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &nctx,
                      const CeedMiOutflowContext &octx, hipStream_t s) {
  switch (kind) {
    case kFaceRiemannOutflow: return bc_diag_var<P, Q, fop::RiemannOutflowJacobian>(pblock, var, fp32, a, octx, s);
    case kFacePressureOutflow: return bc_diag_var<P, Q, fop::PressureOutflowJacobian>(pblock, var, fp32, a, octx, s);
    case kFaceSlip: return bc_diag_var<P, Q, fop::SlipJacobian>(pblock, var, fp32, a, nctx, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

hipError_t launch_face_outflow(int P, int Q, int var, bool pressure, bool jacobian, bool fp32, const VolumeArgs &a, const CeedMiOutflowContext &ctx,
                               hipStream_t s) {
#define CALL_OUT(P_, Q_) (outflow_pq<P_, Q_>(var, pressure, jacobian, fp32, a, ctx, s))
  HONEE_FACE_PQ(CALL_OUT)
#undef CALL_OUT
  return hipErrorInvalidValue;
}

hipError_t launch_face_slip(int P, int Q, int var, bool jacobian, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
#define CALL_SLIP(P_, Q_) (bc_apply_var<P_, Q_, fop::Slip, fop::SlipJacobian>(var, jacobian, fp32, a, ctx, s))
  HONEE_FACE_PQ(CALL_SLIP)
#undef CALL_SLIP
  return hipErrorInvalidValue;
}

hipError_t launch_face_bc_jacobian_diagonal(bool point_block, int P, int Q, int var, int kind, bool fp32, const VolumeArgs &a,
                                            const CeedMiNewtonianContext &nctx, const CeedMiOutflowContext &octx, hipStream_t s) {
#define CALL_BD(P_, Q_) (bc_diag_pq<P_, Q_>(point_block, var, kind, fp32, a, nctx, octx, s))
  HONEE_FACE_PQ(CALL_BD)
#undef CALL_BD
  return hipErrorInvalidValue;
}
