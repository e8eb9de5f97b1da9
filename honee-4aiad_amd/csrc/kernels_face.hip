// Fused kernels of the boundary-face operators on MI355X: the consistent boundary integral and the freestream Riemann
// boundary (instantiations of the face kernel templates in face_kernels.hpp), and k_setup_boundary<Q>, which computes
// the face quadrature data {w detJb, normal, dXdx} from the bilinear coordinate field (SetupBoundary,
// qfunctions/setupgeo.h:100-117).
#include "face_kernels.hpp"

namespace {
template <int P, int Q, bool JAC>
hipError_t face_boundary_integral_var(int var, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx, hipStream_t s) {
  if constexpr (!JAC) {
    switch (var) {
      case 0: return launch_face_op<P, Q, fop::BoundaryIntegral<0>>(a, ctx, s);
      case 1: return launch_face_op<P, Q, fop::BoundaryIntegral<1>>(a, ctx, s);
      default: return launch_face_op<P, Q, fop::BoundaryIntegral<2>>(a, ctx, s);
    }
  } else {
    if (fp32) {
      switch (var) {
        case 0: return launch_face_op<P, Q, fop::BoundaryIntegralJacobian<0, float>>(a, ctx, s);
        case 1: return launch_face_op<P, Q, fop::BoundaryIntegralJacobian<1, float>>(a, ctx, s);
        default: return launch_face_op<P, Q, fop::BoundaryIntegralJacobian<2, float>>(a, ctx, s);
      }
    }
    switch (var) {
      case 0: return launch_face_op<P, Q, fop::BoundaryIntegralJacobian<0, double>>(a, ctx, s);
      case 1: return launch_face_op<P, Q, fop::BoundaryIntegralJacobian<1, double>>(a, ctx, s);
      default: return launch_face_op<P, Q, fop::BoundaryIntegralJacobian<2, double>>(a, ctx, s);
    }
  }
}

template <int P, int Q, bool HLLC, bool JAC>
hipError_t face_freestream_var(int var, bool fp32, const VolumeArgs &a, const CeedMiFreestreamContext &ctx, hipStream_t s) {
  if constexpr (!JAC) {
    switch (var) {
      case 0: return launch_face_op<P, Q, fop::Freestream<0, HLLC>>(a, ctx, s);
      case 1: return launch_face_op<P, Q, fop::Freestream<1, HLLC>>(a, ctx, s);
      default: return launch_face_op<P, Q, fop::Freestream<2, HLLC>>(a, ctx, s);
    }
  } else {
    if (fp32) {
      switch (var) {
        case 0: return launch_face_op<P, Q, fop::FreestreamJacobian<0, HLLC, float>>(a, ctx, s);
        case 1: return launch_face_op<P, Q, fop::FreestreamJacobian<1, HLLC, float>>(a, ctx, s);
        default: return launch_face_op<P, Q, fop::FreestreamJacobian<2, HLLC, float>>(a, ctx, s);
      }
    }
    switch (var) {
      case 0: return launch_face_op<P, Q, fop::FreestreamJacobian<0, HLLC, double>>(a, ctx, s);
      case 1: return launch_face_op<P, Q, fop::FreestreamJacobian<1, HLLC, double>>(a, ctx, s);
      default: return launch_face_op<P, Q, fop::FreestreamJacobian<2, HLLC, double>>(a, ctx, s);
    }
  }
}

template <int P, int Q>
hipError_t face_diag_var(bool pblock, bool freestream, bool hllc, int var, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &nctx,
                         const CeedMiFreestreamContext &fctx, hipStream_t s) {
#define FD_VARS(OPT, CTX)                                                                                              \
  switch (var) {                                                                                                       \
    case 0: return pblock ? launch_face_pblock<P, Q, OPT(0)>(a, CTX, s) : launch_face_diag<P, Q, OPT(0)>(a, CTX, s);  \
    case 1: return pblock ? launch_face_pblock<P, Q, OPT(1)>(a, CTX, s) : launch_face_diag<P, Q, OPT(1)>(a, CTX, s);  \
    default: return pblock ? launch_face_pblock<P, Q, OPT(2)>(a, CTX, s) : launch_face_diag<P, Q, OPT(2)>(a, CTX, s); \
  }
#define BI_D(V) fop::BoundaryIntegralJacobian<V, double>
#define BI_F(V) fop::BoundaryIntegralJacobian<V, float>
#define FS_HLL_D(V) fop::FreestreamJacobian<V, false, double>
#define FS_HLL_F(V) fop::FreestreamJacobian<V, false, float>
#define FS_HLLC_D(V) fop::FreestreamJacobian<V, true, double>
#define FS_HLLC_F(V) fop::FreestreamJacobian<V, true, float>
  if (!freestream) {
    if (fp32) FD_VARS(BI_F, nctx)
    FD_VARS(BI_D, nctx)
  }
  if (hllc) {
    if (fp32) FD_VARS(FS_HLLC_F, fctx)
    FD_VARS(FS_HLLC_D, fctx)
  }
  if (fp32) FD_VARS(FS_HLL_F, fctx)
  FD_VARS(FS_HLL_D, fctx)
#undef FD_VARS
#undef BI_D
#undef BI_F
#undef FS_HLL_D
#undef FS_HLL_F
#undef FS_HLLC_D
#undef FS_HLLC_F
}

}  // namespace


bool face_pq_supported(int P, int Q) {
  return (P == Q && P >= 2 && P <= 5) || (P >= 2 && P <= 4 && Q >= P + 1 && Q <= P + 2);
}

hipError_t launch_setup_boundary(int Q, const SetupArgs &a, hipStream_t stream) {
  const long long n     = (long long)a.nelem * Q * Q;
  const int       block = 256;
  const int       grid  = (int)((n + block - 1) / block);
  if (n == 0) return hipSuccess;
  switch (Q) {
    case 2: hipLaunchKernelGGL((k_setup_boundary<2>), dim3(grid), dim3(block), 0, stream, a); break;
    case 3: hipLaunchKernelGGL((k_setup_boundary<3>), dim3(grid), dim3(block), 0, stream, a); break;
    case 4: hipLaunchKernelGGL((k_setup_boundary<4>), dim3(grid), dim3(block), 0, stream, a); break;
    case 5: hipLaunchKernelGGL((k_setup_boundary<5>), dim3(grid), dim3(block), 0, stream, a); break;
    case 6: hipLaunchKernelGGL((k_setup_boundary<6>), dim3(grid), dim3(block), 0, stream, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_face_boundary_integral(int P, int Q, int var, bool jacobian, bool fp32, const VolumeArgs &a, const CeedMiNewtonianContext &ctx,
                                         hipStream_t s) {
#define CALL_BI(P_, Q_) (jacobian ? face_boundary_integral_var<P_, Q_, true>(var, fp32, a, ctx, s) : face_boundary_integral_var<P_, Q_, false>(var, fp32, a, ctx, s))
  HONEE_FACE_PQ(CALL_BI)
#undef CALL_BI
  return hipErrorInvalidValue;
}

hipError_t launch_face_freestream(int P, int Q, int var, bool hllc, bool jacobian, bool fp32, const VolumeArgs &a, const CeedMiFreestreamContext &ctx,
                                  hipStream_t s) {
#define CALL_FS(P_, Q_)                                                                                                    \
  (hllc ? (jacobian ? face_freestream_var<P_, Q_, true, true>(var, fp32, a, ctx, s) : face_freestream_var<P_, Q_, true, false>(var, fp32, a, ctx, s)) \
        : (jacobian ? face_freestream_var<P_, Q_, false, true>(var, fp32, a, ctx, s) : face_freestream_var<P_, Q_, false, false>(var, fp32, a, ctx, s)))
  HONEE_FACE_PQ(CALL_FS)
#undef CALL_FS
  return hipErrorInvalidValue;
}

hipError_t launch_face_jacobian_diagonal(bool point_block, int P, int Q, int var, bool freestream, bool hllc, bool fp32, const VolumeArgs &a,
                                         const CeedMiNewtonianContext &nctx, const CeedMiFreestreamContext &fctx, hipStream_t s) {
#define CALL_FD(P_, Q_) (face_diag_var<P_, Q_>(point_block, freestream, hllc, var, fp32, a, nctx, fctx, s))
  HONEE_FACE_PQ(CALL_FD)
#undef CALL_FD
  return hipErrorInvalidValue;
}
