// vendor_spgemm.cpp - rocSPARSE's generic SpGEMM on the same operands, as a timing baseline.
//
// Built into its own library (libsmcm_vendor.so); the product path never uses it.  rocSPARSE has
// no integer value types, so values are binary64 (path counts are exact below 2^53) and the
// structure of C is the comparison that is exact.  Times: the nnz and compute stages (buffer and
// output allocations made once, outside the timed region), HIP events on one stream.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

extern "C" {

// status 0 ok; -22 invalid operand (checked before any device work); -5 a HIP or rocSPARSE step
// failed (message in err, cap bytes)
int smcm_vendor_spgemm(int32_t m, int32_t k, int32_t n, const int64_t* a_ptr, const int32_t* a_col,
                       const double* a_val, const int64_t* b_ptr, const int32_t* b_col, const double* b_val,
                       int32_t reps, double* ms_per_product, int64_t* c_nnz, int64_t* c_ptr_out,
                       int32_t* c_col_out, double* c_val_out, char* err, int32_t cap);
}

namespace {

struct Fail {
  std::string msg;
  int code = -5;  // -22: invalid operand (nothing reached the device)
};

#define V_HIP(x)                                                         \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess) throw Fail{std::string(#x) + ": " + hipGetErrorString(e_)}; \
  } while (0)
#define V_SP(x)                                                          \
  do {                                                                   \
    rocsparse_status s_ = (x);                                           \
    if (s_ != rocsparse_status_success) throw Fail{std::string(#x) + ": status " + std::to_string((int)s_)}; \
  } while (0)

// ptr[0] == 0, non-decreasing row pointers, columns present when there are entries and in
// [0, ncols): checked on the host before anything reaches the device
void check_operand(const char* name, int32_t rows, int32_t ncols, const int64_t* ptr, const int32_t* col,
                   const double* val) {
  const std::string nm(name);
  if (!ptr) throw Fail{nm + ": null row pointers", -22};
  if (ptr[0] != 0) throw Fail{nm + ": row pointer 0 must be 0", -22};
  for (int32_t r = 0; r < rows; ++r)
    if (ptr[r + 1] < ptr[r]) throw Fail{nm + ": row pointers decrease at row " + std::to_string(r), -22};
  if (ptr[rows] > 0 && (!col || !val)) throw Fail{nm + ": null column or value array", -22};
  for (int64_t x = 0; x < ptr[rows]; ++x)
    if (col[x] < 0 || col[x] >= ncols) throw Fail{nm + ": column out of range", -22};
}

template <class T>
T* dev_copy(const T* h, size_t n) {
  T* d = nullptr;
  V_HIP(hipMalloc(&d, sizeof(T) * (n ? n : 1)));
  if (n) V_HIP(hipMemcpy(d, h, sizeof(T) * n, hipMemcpyHostToDevice));
  return d;
}

}  // namespace

int smcm_vendor_spgemm(int32_t m, int32_t k, int32_t n, const int64_t* a_ptr, const int32_t* a_col,
                       const double* a_val, const int64_t* b_ptr, const int32_t* b_col, const double* b_val,
                       int32_t reps, double* ms_per_product, int64_t* c_nnz, int64_t* c_ptr_out,
                       int32_t* c_col_out, double* c_val_out, char* err, int32_t cap) {
  std::vector<void*> dev;
  rocsparse_handle h = nullptr;
  rocsparse_spmat_descr A = nullptr, B = nullptr, C = nullptr, D = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  try {
    if (m < 0 || k < 0 || n < 0 || reps < 1) throw Fail{"bad arguments", -22};
    check_operand("A", m, k, a_ptr, a_col, a_val);
    check_operand("B", k, n, b_ptr, b_col, b_val);
    const int64_t annz = a_ptr[m], bnnz = b_ptr[k];
    auto keep = [&](void* p) {
      dev.push_back(p);
      return p;
    };
    void* dap = keep(dev_copy(a_ptr, (size_t)m + 1));
    void* dac = keep(dev_copy(a_col, (size_t)annz));
    void* dav = keep(dev_copy(a_val, (size_t)annz));
    void* dbp = keep(dev_copy(b_ptr, (size_t)k + 1));
    void* dbc = keep(dev_copy(b_col, (size_t)bnnz));
    void* dbv = keep(dev_copy(b_val, (size_t)bnnz));
    V_HIP(hipStreamCreate(&st));
    V_SP(rocsparse_create_handle(&h));
    V_SP(rocsparse_set_stream(h, st));
    const auto i64 = rocsparse_indextype_i64, i32 = rocsparse_indextype_i32;
    const auto f64 = rocsparse_datatype_f64_r;
    V_SP(rocsparse_create_csr_descr(&A, m, k, annz, dap, dac, dav, i64, i32, rocsparse_index_base_zero, f64));
    V_SP(rocsparse_create_csr_descr(&B, k, n, bnnz, dbp, dbc, dbv, i64, i32, rocsparse_index_base_zero, f64));
    void* dcp = keep(nullptr);
    V_HIP(hipMalloc(&dcp, sizeof(int64_t) * ((size_t)m + 1)));
    dev.back() = dcp;
    V_SP(rocsparse_create_csr_descr(&C, m, n, 0, dcp, nullptr, nullptr, i64, i32, rocsparse_index_base_zero, f64));
    // D: an empty m x n matrix (beta = 0)
    std::vector<int64_t> zeros((size_t)m + 1, 0);
    void* ddp = keep(dev_copy(zeros.data(), zeros.size()));
    void* ddc = keep(dev_copy(a_col, 1));
    void* ddv = keep(dev_copy(a_val, 1));
    V_SP(rocsparse_create_csr_descr(&D, m, n, 0, ddp, ddc, ddv, i64, i32, rocsparse_index_base_zero, f64));
    const double alpha = 1.0, beta = 0.0;
    size_t bsize = 0;
    V_SP(rocsparse_spgemm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, A, B, &beta, D, C,
                          f64, rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_buffer_size, &bsize, nullptr));
    void* buf = keep(nullptr);
    V_HIP(hipMalloc(&buf, bsize ? bsize : 1));
    dev.back() = buf;
    V_HIP(hipEventCreate(&e0));
    V_HIP(hipEventCreate(&e1));
    void* dcc = nullptr;
    void* dcv = nullptr;
    int64_t nnz = 0;
    float total_ms = 0;
    for (int32_t r = 0; r <= reps; ++r) {  // r = 0: warm-up (and the output allocation)
      V_HIP(hipEventRecord(e0, st));
      V_SP(rocsparse_spgemm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, A, B, &beta, D, C,
                            f64, rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_nnz, &bsize, buf));
      if (r == 0) {
        int64_t rr, cc;
        V_HIP(hipStreamSynchronize(st));
        V_SP(rocsparse_spmat_get_size(C, &rr, &cc, &nnz));
        V_HIP(hipMalloc(&dcc, sizeof(int32_t) * (nnz ? nnz : 1)));
        dev.push_back(dcc);
        V_HIP(hipMalloc(&dcv, sizeof(double) * (nnz ? nnz : 1)));
        dev.push_back(dcv);
        V_SP(rocsparse_csr_set_pointers(C, dcp, dcc, dcv));
        V_HIP(hipEventRecord(e0, st));
      }
      V_SP(rocsparse_spgemm(h, rocsparse_operation_none, rocsparse_operation_none, &alpha, A, B, &beta, D, C,
                            f64, rocsparse_spgemm_alg_default, rocsparse_spgemm_stage_compute, &bsize, buf));
      V_HIP(hipEventRecord(e1, st));
      V_HIP(hipEventSynchronize(e1));
      float ms = 0;
      V_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (r > 0) total_ms += ms;
    }
    *ms_per_product = total_ms / reps;
    *c_nnz = nnz;
    if (c_ptr_out) V_HIP(hipMemcpy(c_ptr_out, dcp, sizeof(int64_t) * ((size_t)m + 1), hipMemcpyDeviceToHost));
    if (c_col_out && nnz) V_HIP(hipMemcpy(c_col_out, dcc, sizeof(int32_t) * nnz, hipMemcpyDeviceToHost));
    if (c_val_out && nnz) V_HIP(hipMemcpy(c_val_out, dcv, sizeof(double) * nnz, hipMemcpyDeviceToHost));
  } catch (const Fail& f) {
    if (err && cap > 0) {
      std::strncpy(err, f.msg.c_str(), (size_t)cap - 1);
      err[cap - 1] = 0;
    }
    rc = f.code;
  }
  if (A) rocsparse_destroy_spmat_descr(A);
  if (B) rocsparse_destroy_spmat_descr(B);
  if (C) rocsparse_destroy_spmat_descr(C);
  if (D) rocsparse_destroy_spmat_descr(D);
  if (h) rocsparse_destroy_handle(h);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (void* p : dev)
    if (p) (void)hipFree(p);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}
