// iamx_blaslt.hip -- plain dense products through hipBLASLt.
//
// The dense products behind a matmul -- plain FP64 / FP32, the K-stacked
// BF16 product of the split-BF16 form, the INT8 residue products of the
// Ozaki-II form -- are plain library GEMMs once the operands are decoded and
// converted.  The autotuners of those products time this path against the
// hand-written MFMA kernels once per shape class and keep the faster.
//
// Row-major products map onto hipBLASLt's column-major convention as
// C^T = op(B)^T A^T: the "A" operand is B (transposed when B is given as its
// n x k transpose), the "B" operand is A, D is C with its leading dimension.
// The heuristic's first algorithm is kept per shape, with its workspace in a
// slot of its own (IAMX_BLT_WORKSPACE).  Nothing here runs during a graph
// capture: the callers fall back to their own kernels there.
#include <hipblaslt/hipblaslt.h>

#include <mutex>
#include <vector>

#include "iamx_internal.h"

namespace iamx {

namespace {

constexpr size_t kWsBytes = (size_t)64 << 20;

struct BltKey {
  int type, bt, batch;
  int64_t rows, n, k, lda, ldb, ldc, sa, sb, sc;
  bool operator==(const BltKey &o) const {
    return type == o.type && bt == o.bt && batch == o.batch && rows == o.rows && n == o.n && k == o.k &&
           lda == o.lda && ldb == o.ldb && ldc == o.ldc && sa == o.sa && sb == o.sb && sc == o.sc;
  }
};

struct BltPlan {
  BltKey key;
  hipblasLtMatmulDesc_t op;
  hipblasLtMatrixLayout_t la, lb, lc;
  hipblasLtMatmulAlgo_t algo;
  bool ok;
};

hipblasLtHandle_t g_h = nullptr;
int g_state = 0;  // 0 not tried, 1 ready, -1 unavailable
std::mutex g_mu;
std::vector<BltPlan> g_plans;

bool ready() {
  if (g_state) return g_state > 0;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_state) return g_state > 0;
  g_state = hipblasLtCreate(&g_h) == HIPBLAS_STATUS_SUCCESS ? 1 : -1;
  return g_state > 0;
}

void types_of(int type, hipDataType &ab, hipDataType &cd, hipblasComputeType_t &ct, hipDataType &st) {
  switch (type) {
    case BLT_F64: ab = cd = st = HIP_R_64F; ct = HIPBLAS_COMPUTE_64F; break;
    case BLT_F32: ab = cd = st = HIP_R_32F; ct = HIPBLAS_COMPUTE_32F; break;
    case BLT_BF16_F32: ab = HIP_R_16BF; cd = st = HIP_R_32F; ct = HIPBLAS_COMPUTE_32F; break;
    default: ab = HIP_R_8I; cd = st = HIP_R_32I; ct = HIPBLAS_COMPUTE_32I; break;
  }
}

bool set_batch(hipblasLtMatrixLayout_t l, int batch, int64_t stride) {
  if (batch <= 1) return true;
  const int32_t b = batch;
  return hipblasLtMatrixLayoutSetAttribute(l, HIPBLASLT_MATRIX_LAYOUT_BATCH_COUNT, &b, sizeof(b)) ==
             HIPBLAS_STATUS_SUCCESS &&
         hipblasLtMatrixLayoutSetAttribute(l, HIPBLASLT_MATRIX_LAYOUT_STRIDED_BATCH_OFFSET, &stride,
                                           sizeof(stride)) == HIPBLAS_STATUS_SUCCESS;
}

// The plan of one shape: descriptors and the heuristic's first algorithm
// (ok = false when the library offers none).
BltPlan *plan_for(const BltKey &key) {
  for (auto &p : g_plans)
    if (p.key == key) return &p;
  BltPlan p{};
  p.key = key;
  hipDataType ab, cd, st;
  hipblasComputeType_t ct;
  types_of(key.type, ab, cd, ct, st);
  const int32_t ta = key.bt ? HIPBLAS_OP_T : HIPBLAS_OP_N, tb = HIPBLAS_OP_N;
  // stored matrices: "A" = B (n x k, or k x n when transposed), "B" = A (k x rows), D = C (n x rows)
  const int64_t ar = key.bt ? key.k : key.n, ac = key.bt ? key.n : key.k;
  bool ok = hipblasLtMatmulDescCreate(&p.op, ct, st) == HIPBLAS_STATUS_SUCCESS &&
            hipblasLtMatmulDescSetAttribute(p.op, HIPBLASLT_MATMUL_DESC_TRANSA, &ta, sizeof(ta)) ==
                HIPBLAS_STATUS_SUCCESS &&
            hipblasLtMatmulDescSetAttribute(p.op, HIPBLASLT_MATMUL_DESC_TRANSB, &tb, sizeof(tb)) ==
                HIPBLAS_STATUS_SUCCESS &&
            hipblasLtMatrixLayoutCreate(&p.la, ab, ar, ac, key.ldb) == HIPBLAS_STATUS_SUCCESS &&
            hipblasLtMatrixLayoutCreate(&p.lb, ab, key.k, key.rows, key.lda) == HIPBLAS_STATUS_SUCCESS &&
            hipblasLtMatrixLayoutCreate(&p.lc, cd, key.n, key.rows, key.ldc) == HIPBLAS_STATUS_SUCCESS &&
            set_batch(p.la, key.batch, key.sb) && set_batch(p.lb, key.batch, key.sa) &&
            set_batch(p.lc, key.batch, key.sc);
  if (ok) {
    hipblasLtMatmulPreference_t pref;
    ok = hipblasLtMatmulPreferenceCreate(&pref) == HIPBLAS_STATUS_SUCCESS;
    if (ok) {
      const uint64_t ws = kWsBytes;
      hipblasLtMatmulHeuristicResult_t res[1];
      int got = 0;
      ok = hipblasLtMatmulPreferenceSetAttribute(pref, HIPBLASLT_MATMUL_PREF_MAX_WORKSPACE_BYTES, &ws,
                                                 sizeof(ws)) == HIPBLAS_STATUS_SUCCESS &&
           hipblasLtMatmulAlgoGetHeuristic(g_h, p.op, p.la, p.lb, p.lc, p.lc, pref, 1, res, &got) ==
               HIPBLAS_STATUS_SUCCESS &&
           got >= 1 && res[0].workspaceSize <= kWsBytes;
      if (ok) p.algo = res[0].algo;
      (void)hipblasLtMatmulPreferenceDestroy(pref);
    }
  }
  p.ok = ok;
  g_plans.push_back(p);
  return &g_plans.back();
}

}  // namespace

bool blt_usable() { return !capturing() && ready(); }

int blt_gemm(int type, int bt, int64_t rows, int64_t n, int64_t k, const void *A, int64_t lda, const void *B,
             int64_t ldb, void *C, int64_t ldc, int batch, int64_t sa, int64_t sb, int64_t sc, const char *name) {
  if (!blt_usable()) return set_error(IAMX_EINVAL, "hipBLASLt product unavailable");
  if (rows <= 0 || n <= 0 || k <= 0 || batch < 1) return set_error(IAMX_EINVAL, "hipBLASLt product: bad shape");
  void *ws = iamx_workspace(IAMX_BLT_WORKSPACE, kWsBytes);
  if (!ws) return set_error(IAMX_ENOMEM, "hipBLASLt workspace");
  BltKey key{type, bt, batch, rows, n, k, lda, ldb, ldc, batch > 1 ? sa : 0, batch > 1 ? sb : 0, batch > 1 ? sc : 0};
  BltPlan *p = plan_for(key);
  if (!p->ok) return set_error(IAMX_EINVAL, "hipBLASLt offers no algorithm for this product");
  const double one_d = 1.0, zero_d = 0.0;
  const float one_f = 1.0f, zero_f = 0.0f;
  const int32_t one_i = 1, zero_i = 0;
  const void *alpha = type == BLT_F64 ? (const void *)&one_d : type == BLT_I8_I32 ? (const void *)&one_i
                                                                                   : (const void *)&one_f;
  const void *beta = type == BLT_F64 ? (const void *)&zero_d : type == BLT_I8_I32 ? (const void *)&zero_i
                                                                                  : (const void *)&zero_f;
  ++g_launch_seq;
  ProfScope ps(name ? name
               : type == BLT_F64 ? "gemm_f64_hipblaslt"
               : type == BLT_F32 ? "gemm_f32_hipblaslt"
               : type == BLT_BF16_F32 ? "gemm_split_bf16_hipblaslt"
                                      : "gemm_i8_hipblaslt");
  const hipblasStatus_t s = hipblasLtMatmul(g_h, p->op, alpha, B, p->la, A, p->lb, beta, C, p->lc, C, p->lc,
                                            &p->algo, ws, kWsBytes, g_cur);
  if (s != HIPBLAS_STATUS_SUCCESS) return set_error(IAMX_EHIP, "hipblasLtMatmul failed (%d)", (int)s);
  return IAMX_OK;
}

void blt_shutdown() {
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto &p : g_plans) {
    if (p.la) (void)hipblasLtMatrixLayoutDestroy(p.la);
    if (p.lb) (void)hipblasLtMatrixLayoutDestroy(p.lb);
    if (p.lc) (void)hipblasLtMatrixLayoutDestroy(p.lc);
    if (p.op) (void)hipblasLtMatmulDescDestroy(p.op);
  }
  g_plans.clear();
  if (g_h) (void)hipblasLtDestroy(g_h);
  g_h = nullptr;
  g_state = 0;
}

}  // namespace iamx
