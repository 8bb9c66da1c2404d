// C ABI of the SSR hot-path library: runtime helpers (device memory, copies,
// stream, events) and the host-pointer entry points that keep the signatures
// of the reference's MEX files.  Declarations: include/ssr_hip.h.
#include <atomic>
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "ssr_common.h"

static thread_local char g_err[512] = "";
static hipStream_t g_stream = nullptr;

void ssr_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// A thread may divert the library's launches to another stream for the
// duration of a graph capture (the legacy default stream cannot be captured).
static thread_local hipStream_t t_capture = nullptr;
static thread_local bool t_capturing = false;

// A thread may also run on its own stream (concurrent independent problems,
// one host thread each): ssr_set_thread_stream.
static thread_local hipStream_t t_stream = nullptr;
static thread_local bool t_has_stream = false;

hipStream_t ssr_stream(void) { return t_capturing ? t_capture : (t_has_stream ? t_stream : g_stream); }

void ssr_capture_stream_push(hipStream_t s)
{
    t_capture = s;
    t_capturing = true;
}

void ssr_capture_stream_pop(void) { t_capturing = false; }

// true once any host thread has run the library on a stream of its own
// (concurrent instances): graph capture is then left alone
static std::atomic<bool> g_thread_streams{false};
bool ssr_thread_streams_used(void) { return g_thread_streams.load(); }

int ssr_cu_count(void)
{
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (cached[dev] <= 0) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0) c = 256;
        cached[dev] = c;
    }
    return cached[dev];
}

extern "C" {

const char *ssr_last_error(void) { return g_err; }

int ssr_version(void) { return 1; }

int ssr_device_count(int *count)
{
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    *count = c;
    return SSR_OK;
}

int ssr_set_stream(void *stream)
{
    g_stream = (hipStream_t)stream;
    return SSR_OK;
}

void *ssr_get_stream(void) { return (void *)ssr_stream(); }

// Launch everything this host thread issues on `stream` (NULL: back to the
// library stream).  Each thread's problems then proceed concurrently with
// other threads' on the device.
int ssr_set_thread_stream(void *stream)
{
    t_stream = (hipStream_t)stream;
    t_has_stream = stream != nullptr;
    if (stream) g_thread_streams.store(true);
    return SSR_OK;
}

// A new stream at the default priority that does not synchronise with the
// legacy default stream.
int ssr_stream_create(void **out)
{
    SSR_REQUIRE(out, "stream_create: null out pointer");
    hipStream_t st = nullptr;
    SSR_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    *out = (void *)st;
    return SSR_OK;
}

int ssr_stream_destroy(void *stream)
{
    if (stream) SSR_HIP_CHECK(hipStreamDestroy((hipStream_t)stream));
    return SSR_OK;
}

int ssr_dmalloc(void **ptr, size_t bytes)
{
    SSR_REQUIRE(ptr, "dmalloc: null out pointer");
    *ptr = nullptr;
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        ssr_set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return SSR_ERR_ALLOC;
    }
    return SSR_OK;
}

int ssr_dfree(void *ptr)
{
    if (ptr) SSR_HIP_CHECK(hipFree(ptr));
    return SSR_OK;
}

int ssr_htod(void *dst, const void *src, size_t bytes)
{
    if (bytes) SSR_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ssr_stream()));
    SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
    return SSR_OK;
}

int ssr_dtoh(void *dst, const void *src, size_t bytes)
{
    if (bytes) SSR_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ssr_stream()));
    SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
    return SSR_OK;
}

int ssr_dtod(void *dst, const void *src, size_t bytes)
{
    if (bytes) SSR_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ssr_stream()));
    return SSR_OK;
}

int ssr_dzero(void *dst, size_t bytes)
{
    if (bytes) SSR_HIP_CHECK(hipMemsetAsync(dst, 0, bytes, ssr_stream()));
    return SSR_OK;
}

int ssr_synchronize(void)
{
    SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
    return SSR_OK;
}

int ssr_event_create(void **ev)
{
    hipEvent_t e;
    SSR_HIP_CHECK(hipEventCreate(&e));
    *ev = (void *)e;
    return SSR_OK;
}

int ssr_event_destroy(void *ev)
{
    if (ev) SSR_HIP_CHECK(hipEventDestroy((hipEvent_t)ev));
    return SSR_OK;
}

int ssr_event_record(void *ev)
{
    SSR_HIP_CHECK(hipEventRecord((hipEvent_t)ev, ssr_stream()));
    return SSR_OK;
}

int ssr_event_elapsed_ms(void *start, void *stop, float *ms)
{
    SSR_HIP_CHECK(hipEventSynchronize((hipEvent_t)stop));
    SSR_HIP_CHECK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return SSR_OK;
}

int ssr_dev_constitutive_3D(int64_t n, const double *E, const double *c_bar, const double *sin_phi,
                            const double *shear, const double *bulk, const double *lame, double *S,
                            double *DS, int8_t *branch);
int ssr_dev_constitutive_2D(int64_t n, const double *E, const double *c_bar, const double *sin_phi,
                            const double *shear, const double *bulk, const double *lame, double *S,
                            double *DS, int8_t *branch);

}  // extern "C"

// Host-pointer path shared by the four constitutive MEX-signature entry points.
namespace {
struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t bytes) { return ssr_dmalloc(&p, bytes); }
    double *d() { return (double *)p; }
};

int constitutive_host(int dim, int64_t n, const double *E, const double *c_bar, const double *sin_phi,
                      const double *shear, const double *bulk, const double *lame, double *S,
                      double *DS)
{
    SSR_REQUIRE(n >= 0, "n_int must be >= 0");
    if (n == 0) return SSR_OK;
    SSR_REQUIRE(E && c_bar && sin_phi && shear && bulk && lame && S, "null argument");
    const int ne = dim == 3 ? 6 : 3, nsd = dim == 3 ? 6 : 4, nds = dim == 3 ? 36 : 9;
    DevBuf dE, dm, dS, dD;
    int rc;
    if ((rc = dE.alloc(sizeof(double) * ne * n))) return rc;
    if ((rc = dm.alloc(sizeof(double) * 5 * n))) return rc;
    if ((rc = dS.alloc(sizeof(double) * nsd * n))) return rc;
    if (DS && (rc = dD.alloc(sizeof(double) * nds * n))) return rc;
    if ((rc = ssr_htod(dE.p, E, sizeof(double) * ne * n))) return rc;
    const double *src[5] = {c_bar, sin_phi, shear, bulk, lame};
    for (int k = 0; k < 5; ++k)
        if ((rc = ssr_htod(dm.d() + k * n, src[k], sizeof(double) * n))) return rc;
    if (dim == 3)
        rc = ssr_dev_constitutive_3D(n, dE.d(), dm.d(), dm.d() + n, dm.d() + 2 * n, dm.d() + 3 * n,
                                     dm.d() + 4 * n, dS.d(), DS ? dD.d() : nullptr, nullptr);
    else
        rc = ssr_dev_constitutive_2D(n, dE.d(), dm.d(), dm.d() + n, dm.d() + 2 * n, dm.d() + 3 * n,
                                     dm.d() + 4 * n, dS.d(), DS ? dD.d() : nullptr, nullptr);
    if (rc) return rc;
    if ((rc = ssr_dtoh(S, dS.p, sizeof(double) * nsd * n))) return rc;
    if (DS && (rc = ssr_dtoh(DS, dD.p, sizeof(double) * nds * n))) return rc;
    return SSR_OK;
}
}  // namespace

extern "C" {

int ssr_constitutive_problem_3D_SDS(int64_t n_int, const double *E, const double *c_bar,
                                    const double *sin_phi, const double *shear, const double *bulk,
                                    const double *lame, double *S, double *DS)
{
    SSR_REQUIRE(DS, "constitutive_problem_3D_SDS: DS output required");
    return constitutive_host(3, n_int, E, c_bar, sin_phi, shear, bulk, lame, S, DS);
}

int ssr_constitutive_problem_3D_S(int64_t n_int, const double *E, const double *c_bar,
                                  const double *sin_phi, const double *shear, const double *bulk,
                                  const double *lame, double *S)
{
    return constitutive_host(3, n_int, E, c_bar, sin_phi, shear, bulk, lame, S, nullptr);
}

int ssr_constitutive_problem_2D_SDS(int64_t n_int, int64_t rows_E, const double *E,
                                    const double *c_bar, const double *sin_phi, const double *shear,
                                    const double *bulk, const double *lame, double *S, double *DS)
{
    SSR_REQUIRE(rows_E == 3, "constitutive_problem_2D_SDS_mex: E must have 3 rows.");
    SSR_REQUIRE(DS, "constitutive_problem_2D_SDS: DS output required");
    return constitutive_host(2, n_int, E, c_bar, sin_phi, shear, bulk, lame, S, DS);
}

int ssr_constitutive_problem_2D_S(int64_t n_int, int64_t rows_E, const double *E,
                                  const double *c_bar, const double *sin_phi, const double *shear,
                                  const double *bulk, const double *lame, double *S)
{
    SSR_REQUIRE(rows_E == 3, "constitutive_problem_2D_S_mex: E must have 3 rows.");
    return constitutive_host(2, n_int, E, c_bar, sin_phi, shear, bulk, lame, S, nullptr);
}

}  // extern "C"

extern "C" int ssr_dev_assemble_K_tangent_vals(int dim, int64_t n_p, int64_t n_int, int64_t n_e, int n_q,
                                               const double *DPhi1, const double *DPhi2,
                                               const double *DPhi3, const double *DS,
                                               const double *WEIGHT, const int64_t *smap, int64_t nnz,
                                               double *V);

// Host-pointer drop-in for assemble_K_tangent_vals{,_2D} (MEX boundary).
extern "C" int ssr_assemble_K_tangent_vals(int dim, int64_t n_p, int64_t n_int, int64_t n_e, int n_q,
                                           const double *DPhi1, const double *DPhi2, const double *DPhi3,
                                           const double *DS, const double *WEIGHT,
                                           const int64_t *scatter_map, int64_t nnz_out, double *V)
{
    SSR_REQUIRE(dim == 2 || dim == 3, "assemble_K_tangent_vals: dim must be 2 or 3");
    SSR_REQUIRE(scatter_map, "assemble_K_tangent_vals: scatter_map must be int64.");
    SSR_REQUIRE(n_e * n_q == n_int, "assemble_K_tangent_vals: n_int != n_e*n_q");
    if (nnz_out == 0) return SSR_OK;
    SSR_REQUIRE(V && DPhi1 && DPhi2 && DS && WEIGHT && (dim == 2 || DPhi3), "assemble_K_tangent_vals: null argument");
    const int ns = dim == 3 ? 6 : 3, nld = dim * (int)n_p;
    DevBuf dP, dD, dW, dM, dV;
    int rc;
    if ((rc = dP.alloc(sizeof(double) * dim * n_p * n_int))) return rc;
    if ((rc = dD.alloc(sizeof(double) * ns * ns * n_int))) return rc;
    if ((rc = dW.alloc(sizeof(double) * n_int))) return rc;
    if ((rc = dM.alloc(sizeof(int64_t) * nld * nld * n_e))) return rc;
    if ((rc = dV.alloc(sizeof(double) * nnz_out))) return rc;
    const double *src[3] = {DPhi1, DPhi2, DPhi3};
    for (int k = 0; k < dim; ++k)
        if ((rc = ssr_htod(dP.d() + k * n_p * n_int, src[k], sizeof(double) * n_p * n_int))) return rc;
    if ((rc = ssr_htod(dD.p, DS, sizeof(double) * ns * ns * n_int))) return rc;
    if ((rc = ssr_htod(dW.p, WEIGHT, sizeof(double) * n_int))) return rc;
    if ((rc = ssr_htod(dM.p, scatter_map, sizeof(int64_t) * nld * nld * n_e))) return rc;
    rc = ssr_dev_assemble_K_tangent_vals(dim, n_p, n_int, n_e, n_q, dP.d(), dP.d() + n_p * n_int,
                                         dim == 3 ? dP.d() + 2 * n_p * n_int : nullptr, dD.d(), dW.d(),
                                         (const int64_t *)dM.p, nnz_out, dV.d());
    if (rc) return rc;
    return ssr_dtoh(V, dV.p, sizeof(double) * nnz_out);
}

static int g_tuning[SSR_TUNE_COUNT] = {11, 0, 0, 0, 1, 0, 0, 0, 0, 1, 1, 0, 1, 1, 0, 1, 1, 1, 1};

int ssr_tuning(int key) { return (key >= 0 && key < SSR_TUNE_COUNT) ? g_tuning[key] : 0; }

// Set a kernel-variant / launch-shape knob (see SSR_TUNE_* in ssr_common.h).
extern "C" int ssr_set_tuning(int key, int value)
{
    SSR_REQUIRE(key >= 0 && key < SSR_TUNE_COUNT, "set_tuning: unknown key %d", key);
    if (key == SSR_TUNE_SPMV_GROUP)
        SSR_REQUIRE(value == 0 || value == 4 || value == 8 || value == 16 || value == 32 || value == 64,
                    "set_tuning: SpMV group must be 0 (by dimension), 4, 8, 16, 32 or 64");
    if (key == SSR_TUNE_SPMM_GROUP)
        SSR_REQUIRE(value == 0 || value == 4 || value == 8 || value == 16 || value == 32,
                    "set_tuning: multi-vector SpMV group must be 0 (by dimension), 4, 8, 16 or 32");
    if (key == SSR_TUNE_SPMV_UNROLL)
        SSR_REQUIRE(value >= 0 && value <= 2, "set_tuning: SpMV blocks per lane must be 0 (by dimension), 1 or 2");
    if (key == SSR_TUNE_ELEMENT_KERNEL)
        SSR_REQUIRE(value >= 0 && value <= 2, "set_tuning: element kernel must be 0, 1 or 2");
    // Replay of the V-cycle from a captured graph is switched off: one profiled
    // FP32 continuation ended in a host SIGSEGV inside Amg::apply_graph<float>
    // whose cause is not known, so the form cannot be selected.
    SSR_REQUIRE(key != SSR_TUNE_AMG_GRAPH || value == 0,
                "set_tuning: V-cycle graph replay is disabled (a host fault in apply_graph has no known cause)");
    g_tuning[key] = value;
    return SSR_OK;
}

extern "C" int ssr_get_tuning(int key, int *value)
{
    SSR_REQUIRE(key >= 0 && key < SSR_TUNE_COUNT && value, "get_tuning: unknown key %d", key);
    *value = g_tuning[key];
    return SSR_OK;
}
