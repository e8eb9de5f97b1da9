// Internal object model of the MI355X libCEED-compatible library.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "ceed.h"
#include "ceed/backend.h"
#include "honee_mi355x.h"
#include "kernels.hpp"

// Every object starts with the Ceed it belongs to, so that a cast to CeedObject reads it (CeedObjectReturnCeed); a
// Ceed starts with itself.
struct Ceed_private {
  Ceed             self     = nullptr;
  int              refcount = 1;
  std::string      resource;
  int              device   = 0;
  hipStream_t      stream   = nullptr;  // null stream unless CeedSetStream
  bool             jac_fp32 = false;
  int              scatter  = 1;  // ScatterMode: 0 atomic, 1 two-pass (deterministic, the default)
  bool             qdata_recompute = false;  // volume kernels recompute qdata from a GeoRecord (qdata=recompute)
  double          *d_work   = nullptr;  // E-vector workspace of the two-pass scatter
  size_t           work_len = 0;
  CeedErrorHandler handler  = nullptr;
  std::string      err_msg;
  // kernel timing
  bool                                     timing = false;
  std::map<std::string, std::pair<int, double>> kernel_times;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending_events;
  std::vector<hipEvent_t>                  event_pool;
  hipEvent_t                               region_start = nullptr;  // open CeedMiTimeRegionBegin
  double                                  *d_scratch = nullptr;  // reduction scratch (norms)
  double                                  *h_scratch = nullptr;
  int                                      scratch_len = 0;
  double *d_fixed = nullptr;  // this device's copy of the fixed E-vector scratch (CeedMiOperatorApplyNoAlloc), or null
};

// True on this thread while CeedMiOperatorApplyNoAlloc runs: every helper that would allocate returns an error instead
extern thread_local bool g_no_alloc;

// Geometry of a qdata vector written by the Setup operator (setupgeo.h:42-56): the element vertex coordinates and the
// 1D tables of the trilinear coordinate basis it was computed from.  Volume kernels that read this vector as qdata
// recompute {w detJ, dXdx} from it instead of streaming 10 values per quadrature point (identical values: the setup
// kernel evaluates the same device function).  Any write to the vector drops the record.
struct GeoRecord {
  double *d_xe  = nullptr;  // [nelem][3][8] vertex coordinates, lexicographic vertices
  double *d_tab = nullptr;  // Bx [Q][2], Dx [Q][2], w [Q]
  int     nelem = 0, Q = 0;
  QDataLayout qd{};         // strided layout the values were written in
  ~GeoRecord();
};

struct CeedVector_private {
  Ceed     ceed;
  int      refcount = 1;
  CeedSize length   = 0;
  double  *h_array  = nullptr;
  double  *d_array  = nullptr;
  bool     h_owned  = false, d_owned = false;
  bool     h_valid  = false, d_valid = false;
  int      num_readers = 0;
  bool     writing     = false;
  float   *d_shadow32  = nullptr;  // FP32 mirror (backend-internal; jac_data in FP32 Jacobian mode)
  CeedSize shadow_len  = 0;
  std::shared_ptr<GeoRecord> geo;  // set by the Setup operator on its qdata output
};

struct CeedElemRestriction_private {
  Ceed                 ceed;
  int                  refcount = 1;
  CeedInt              num_elem = 0, elem_size = 0, num_comp = 0, comp_stride = 1;
  CeedSize             l_size   = 0;
  bool                 strided  = false;
  bool                 backend_strides = false;
  CeedInt              strides[3] = {0, 0, 0};
  std::vector<CeedInt> h_offsets;
  int                 *d_offsets = nullptr;
  // node-sorted E-vector map of the two-pass scatter: element node i = e * elem_size + j writes its contribution to
  // slot t_perm[i]; the slots of L-vector row k (offset t_offset[k]) are t_rowptr[k]..t_rowptr[k+1]-1, ordered by i
  int                  t_nrows      = 0;
  bool                 t_covers_all = false;  // the rows cover every L-vector entry
  int                 *d_t_rowptr = nullptr, *d_t_perm = nullptr, *d_t_offset = nullptr;
  int                 *d_t_index = nullptr;  // element node of each slot (inverse of t_perm)
};

struct CeedBasis_private {
  Ceed                ceed;
  int                 refcount = 1;
  bool                tensor   = true;
  CeedInt             dim = 0, num_comp = 0, P_1d = 0, Q_1d = 0, num_nodes = 0, num_qpts = 0;
  CeedElemTopology    topo = CEED_TOPOLOGY_HEX;
  std::vector<double> interp_1d, grad_1d, q_ref_1d, q_weight_1d, collo_grad_1d;
  std::vector<double> interp, grad, q_ref, q_weight;  // non-tensor bases
  double             *d_interp_1d = nullptr, *d_grad_1d = nullptr, *d_collo = nullptr, *d_weight_1d = nullptr;
  double             *d_prod = nullptr;  // [3][Q][P]: B.B, G.B, G.G
  bool                is_projection = false;
};

struct QFField {
  std::string  name;
  CeedInt      size;
  CeedEvalMode emode;
};

enum class QfKind {
  Unknown,
  Identity,
  Setup,
  IFunction,
  IJacobian,
  RHSFunction,
  MassNewtonian,
  Mass5,
  MonitorKE,
  SetupBoundary,
  BoundaryIntegral,
  BoundaryIntegralJacobian,
  Freestream,
  FreestreamJacobian,
  RiemannOutflow,
  RiemannOutflowJacobian,
  PressureOutflow,
  PressureOutflowJacobian,
  Slip,
  SlipJacobian,
  VelocityGradientRHS,
  AnisotropyRHS,
  MassN,
  InverseMultiplicity,
  SgsNodal,
  NodalSgsResidual,
  DiffFilterRHS,
  DiffFilterLHS,
  AnisotropyCollocate,
  SgsTrainingNodal,
};

struct CeedQFunction_private {
  Ceed                 ceed;
  int                  refcount = 1;
  CeedInt              vec_length = 1;
  CeedQFunctionUser    user       = nullptr;
  std::string          source, name;
  std::vector<QFField> inputs, outputs;
  CeedQFunctionContext ctx = nullptr;
  CeedSize             user_flops = 0;
  QfKind               kind       = QfKind::Unknown;
  int                  state_var  = 0;
  bool                 hllc       = false;  // Freestream*: HLLC (true) or HLL Riemann flux
  int                  ncomp      = 0;      // Mass_N, DifferentialFilter_LHS_N: N
};

struct CtxField {
  std::string          name, description;
  size_t               offset, num_values;
  CeedContextFieldType type;
};

struct CeedContextFieldLabel_private {
  std::string                       name;
  CeedContextFieldType              type;
  size_t                            num_values;
  std::vector<std::pair<CeedQFunctionContext, size_t>> targets;  // (context, byte offset)
};

struct CeedQFunctionContext_private {
  Ceed                                  ceed;
  int                                   refcount = 1;
  size_t                                size     = 0;
  void                                 *data     = nullptr;
  bool                                  owned    = false;
  CeedQFunctionContextDataDestroyUser   destroy  = nullptr;
  std::vector<CtxField>                 fields;
  std::vector<std::unique_ptr<CeedContextFieldLabel_private>> labels;
};

struct CeedOperatorField_private {
  std::string         name;
  CeedElemRestriction rstr  = nullptr;
  CeedBasis           basis = nullptr;
  CeedVector          vec   = nullptr;
};

struct OpPlan;  // backend binding of a non-composite operator

// CeedOperatorGetOperatorAssemblyData: the distinct active bases of the operator's input and output fields
struct CeedOperatorAssemblyData_private {
  std::vector<CeedBasis> bases_in, bases_out;
};

struct CeedOperator_private {
  Ceed                                     ceed;
  int                                      refcount = 1;
  bool                                     composite = false;
  CeedQFunction                            qf = nullptr;
  std::vector<CeedOperatorField_private *> in_fields, out_fields;
  std::vector<CeedOperator>                subs;
  std::string                              name;
  CeedInt                                  num_tabs = 0;
  std::vector<std::unique_ptr<CeedContextFieldLabel_private>> labels;
  OpPlan                                  *plan = nullptr;
  CeedOperatorAssemblyData_private        *asm_data = nullptr;
};

// ---- helpers shared by the translation units ----
int  ceed_set_error(Ceed ceed, const char *file, int line, const char *func, int ecode, const char *fmt, ...);
#define MI_ERR(ceed, code, ...) ceed_set_error((ceed), __FILE__, __LINE__, __func__, (code), __VA_ARGS__)
#define MI_CALL(expr)                \
  do {                               \
    int ierr_ = (expr);              \
    if (ierr_) return ierr_;         \
  } while (0)
#define MI_HIP(ceed, expr)                                                                                              \
  do {                                                                                                                 \
    hipError_t herr_ = (expr);                                                                                         \
    if (herr_ != hipSuccess) return MI_ERR((ceed), CEED_ERROR_BACKEND, "HIP error %s: %s", hipGetErrorName(herr_), #expr); \
  } while (0)

int vec_device_read(CeedVector v, const double **d);
int rstr_transpose_map(CeedElemRestriction r);
// Element-node slots grouped by offset, in increasing offset order, each group listing its slots in increasing order:
// rows r = 0..toff.size()-1 hold slots index[rowptr[r] .. rowptr[r+1]) of offset toff[r].  O(n + max offset).
void group_by_offset(const std::vector<CeedInt> &offsets, std::vector<int> &rowptr, std::vector<int> &index, std::vector<int> &toff);
int ceed_workspace(Ceed ceed, size_t len, double **d);
int vec_device_write(CeedVector v, double **d, bool keep_contents);
int vec_device_shadow32(CeedVector v, float **d, bool create);
int rstr_device_offsets(CeedElemRestriction r, const int **d);
int basis_device_tables(CeedBasis b);
int op_plan_apply(CeedOperator op, CeedVector in, CeedVector out, bool overwrite);
bool op_plan_overwrites(CeedOperator op, CeedVector out);
int op_plan_diagonal(CeedOperator op, CeedVector diag, bool point_block);
void op_plan_destroy(OpPlan *plan);
int timing_begin(Ceed ceed, const char *name, hipEvent_t *start);
int timing_end(Ceed ceed, const char *name, hipEvent_t start);
