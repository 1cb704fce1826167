// A stand-in for the HIP runtime and the kernel library (libsg_hip.so) so
// that the host layer (synaptogenesis_amd/csrc/host) can run on a machine
// without a GPU, under AddressSanitizer: device memory is host memory,
// copies are memcpy, kernels do nothing, streams, events and graphs are
// bookkeeping.  Results are meaningless; what is checked is the host code's
// memory use and the ordering calls it makes.  A wait on an event that was
// never recorded (a scheduler bug: it would order nothing) is counted and
// reported at exit.  Test infrastructure only.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "sg_hip.h"

namespace {
struct StubEvent { bool recorded = false; };
struct StubStream { int capturing = 0; };
std::atomic<long> g_bad_waits(0), g_waits(0), g_records(0), g_launches(0);
struct Report {
    ~Report() {
        if (getenv("HIP_STUB_REPORT"))
            fprintf(stderr, "hip stub: %ld launches, %ld event records, "
                    "%ld waits, %ld waits on unrecorded events\n",
                    g_launches.load(), g_records.load(), g_waits.load(),
                    g_bad_waits.load());
    }
} g_report;
}  // namespace

extern "C" {

long hip_stub_bad_waits(void) { return g_bad_waits.load(); }

hipError_t hipMalloc(void **p, size_t n) {
    *p = calloc(1, n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void *p) { free(p); return hipSuccess; }
hipError_t hipHostMalloc(void **p, size_t n, unsigned int) {
    return hipMalloc(p, n);
}
hipError_t hipHostFree(void *p) { free(p); return hipSuccess; }
hipError_t hipHostGetDevicePointer(void **dp, void *p, unsigned int) {
    *dp = p;
    return hipSuccess;
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
    if (n) memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k,
                          hipStream_t) {
    ++g_launches;
    return hipMemcpy(d, s, n, k);
}
hipError_t hipMemset(void *d, int v, size_t n) {
    memset(d, v, n);
    return hipSuccess;
}
hipError_t hipDeviceSynchronize(void) { return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
hipError_t hipGetLastError(void) { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "hip stub error"; }
hipError_t hipStreamCreate(hipStream_t *s) {
    *s = reinterpret_cast<hipStream_t>(new StubStream());
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    delete reinterpret_cast<StubStream *>(s);
    return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e) {
    *e = reinterpret_cast<hipEvent_t>(new StubEvent());
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) {
    return hipEventCreate(e);
}
hipError_t hipEventDestroy(hipEvent_t e) {
    delete reinterpret_cast<StubEvent *>(e);
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
    ++g_records;
    reinterpret_cast<StubEvent *>(e)->recorded = true;
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) {
    *ms = 0.001f;
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t e, unsigned) {
    ++g_waits;
    if (!reinterpret_cast<StubEvent *>(e)->recorded) ++g_bad_waits;
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    reinterpret_cast<StubStream *>(s)->capturing = 1;
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
    reinterpret_cast<StubStream *>(s)->capturing = 0;
    *g = reinterpret_cast<hipGraph_t>(malloc(8));
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *x, hipGraph_t, hipGraphNode_t *,
                               char *, size_t) {
    *x = reinterpret_cast<hipGraphExec_t>(malloc(8));
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) {
    ++g_launches;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) { free(g); return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t x) {
    free(x);
    return hipSuccess;
}

// ---- kernel library: every launch does nothing -----------------------------
#define SG_STUB(name, ...) int name(__VA_ARGS__) { ++g_launches; return 0; }
SG_STUB(sg_hip_activate_base, const sg_conn *, void *)
SG_STUB(sg_hip_activate_second_order, const sg_conn *, float *, void *)
SG_STUB(sg_hip_tss_step, int32_t *, const int32_t *, const uint8_t *,
        const uint8_t *, const uint32_t *, int, void *)
SG_STUB(sg_hip_iz_activate, const sg_conn *, const sg_iz_syn *,
        const sg_iz_dest *, int, void *)
SG_STUB(sg_hip_iz_update, const sg_conn *, const sg_iz_syn *,
        const sg_iz_dest *, int, const float *, const float *, void *)
SG_STUB(sg_hip_iz_update_activate, const sg_conn *, const sg_iz_syn *,
        const sg_iz_dest *, int, int, const float *, const float *, void *)
SG_STUB(sg_hip_iz_source_update, float *, const int32_t *, int, int, int,
        void *)
SG_STUB(sg_hip_iz_neuron, const sg_iz_layer *, void *)
SG_STUB(sg_hip_pack_plane, const uint32_t *, int, int, uint64_t *, void *)
SG_STUB(sg_hip_gol_neuron, float *, uint32_t *, int, int, int, int, int, int,
        void *)
SG_STUB(sg_hip_rate_neuron, float *, float *, int, int, void *)
SG_STUB(sg_hip_relay_neuron, float *, float *, int, int, int, void *)
SG_STUB(sg_hip_gol_step, const uint32_t *, uint32_t *, float *, int, int, int,
        const float *, int, int, int, int, void *)
SG_STUB(sg_hip_gol_steps, const uint32_t *, uint32_t *, float *, int, int,
        int, const float *, int, int, int, int, int, void *)
int sg_hip_gol_max_steps(const float *k, int rows, int cols) {
    for (int q = 0; q < 9; ++q)
        if (k[q] != 0.0f && k[q] != 1.0f) return 4;
    (void)rows;
    (void)cols;
    return 8;
}
SG_STUB(sg_hip_init_random, float *, int, int, float, float, float, int,
        uint64_t, uint64_t, const uint64_t *, void *)
SG_STUB(sg_hip_add_u64, uint64_t *, uint64_t, void *)
SG_STUB(sg_hip_set_f32, float *, float, long, int, void *)
SG_STUB(sg_hip_transpose32, const void *, void *, int, int, void *)
SG_STUB(sg_hip_tile32, const void *, void *, int, int, int, int, int, void *)
SG_STUB(sg_hip_record32, const void *, void *, int, int, int, int, int, void *)
SG_STUB(sg_hip_copy32, const void *, void *, long, void *)
SG_STUB(sg_hip_dendrite, float *, float *, int, int, float, void *)

int sg_hip_iz_neurons(const sg_iz_layer *layers, int n, void *) {
    // the layers' arrays are read as the kernel would find them
    for (int i = 0; i < n; ++i) {
        volatile float probe = layers[i].voltage[layers[i].size - 1];
        (void)probe;
    }
    ++g_launches;
    return 0;
}
// the engine's environment I/O: copied, so that the sanitizer sees both
// ranges
int sg_hip_copy_u32(const void *src, void *dst, long n, void *) {
    ++g_launches;
    if (n > 0) memmove(dst, src, (size_t)n * 4);
    return 0;
}
int sg_hip_device_count(void) { return 1; }
const char *sg_hip_error_string(int) { return "hip stub error"; }

}  // extern "C"
