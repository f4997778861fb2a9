// Allocation-free graphical QP solve (qk_rgw_qp_graph_noalloc): the Newton
// iteration of GW/RGW_QP_graph.f90:50-92 on device-resident inputs, for
// callers that cannot allocate device memory (code under hipGraph capture,
// serving loops, harnesses that forbid hipMalloc).  It runs the same kernels
// as qk_rgw_qp_graph (k_sigma_partial, k_qp_update, k_qp_init, k_qp_finish)
// with their Newton state and partial sums in a library-owned static device
// array, and a fixed QP_MAXIT = 64 sweeps whose finished orbitals are masked
// on the device (the kernels skip them), so no host synchronisation or
// allocation happens inside: the call only enqueues work on the caller's
// stream.  Every call on one GPU shares that GPU's array; calls that overlap
// on one GPU race.
#include <mutex>

#include "qk_kernels.h"

namespace qk {
namespace {

constexpr int QP_MAXIT = 64;        // GW/RGW_QP_graph.f90:31
constexpr double QP_THRESH = 1e-6;  // GW/RGW_QP_graph.f90:32
constexpr size_t SCRATCH_BYTES = size_t(4) << 20;

// The scratch buffer: 4 MiB per GPU, held until the process ends.  Its name
// is a host-side stand-in; host code uses the device address obtained with
// hipGetSymbolAddress (scratch_ptr) and never the name itself.
static __device__ __align__(256) double scratch[SCRATCH_BYTES / sizeof(double)];

constexpr int MAX_DEVICES = 64;
std::mutex g_mtx;
void* g_ptr[MAX_DEVICES] = {};

thread_local std::string g_err;

// Device address of the current device's copy of scratch (looked up once per
// device, with that device current).
void* scratch_ptr() {
    int dev = 0;
    QK_HIP(hipGetDevice(&dev));
    QK_CHECK(dev >= 0 && dev < MAX_DEVICES, QK_E_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mtx);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        QK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        QK_CHECK(p != nullptr, QK_E_HIP, "hipGetSymbolAddress returned no address");
        g_ptr[dev] = p;
    }
    return g_ptr[dev];
}

// Bytes of scratch one call needs: partial sums 3 n nchunk, w, f, df (n
// each) as doubles, then nit, done (n each) and the active count as ints.
size_t need_bytes(int n, int nchunk) {
    return sizeof(double) * (3 * (size_t)n * nchunk + 3 * (size_t)n) + sizeof(int) * (2 * (size_t)n + 1);
}

}  // namespace
}  // namespace qk

extern "C" {

int qk_rgw_qp_graph_noalloc(int doSRG, double eta, double flow, int nOrb, int nC, int nO, int nV, int nR, int nS,
                            const double* eHF, const double* Om, const double* rhoT, const double* eGWlin,
                            const double* eOld, double* eGW, double* Z, void* stream) {
    using namespace qk;
    try {
        QK_CHECK(eHF && Om && rhoT && eGWlin && eOld && eGW && Z, QK_E_ARG, "null device pointer");
        QK_CHECK(nOrb > 0 && nC >= 0 && nR >= 0 && nO > nC && nO < nOrb - nR && nV == nOrb - nO, QK_E_ARG,
                 "need 0 <= nC < nO < nOrb - nR and nV = nOrb - nO");
        QK_CHECK((long long)nS == (long long)(nO - nC) * (nV - nR), QK_E_ARG, "need nS = (nO-nC)*(nV-nR)");
        const int nchunk = sigma_chunks(nS, nOrb - nR - nC);
        QK_CHECK(need_bytes(nOrb, nchunk) <= SCRATCH_BYTES, QK_E_ARG,
                 "qk_rgw_qp_graph_noalloc: nOrb = " + std::to_string(nOrb) + ", nS = " + std::to_string(nS) +
                     " need " + std::to_string(need_bytes(nOrb, nchunk)) + " bytes of scratch, more than its " +
                     std::to_string(SCRATCH_BYTES));
        hipStream_t s = static_cast<hipStream_t>(stream);
        double* part = static_cast<double*>(scratch_ptr());
        double* w = part + 3 * (size_t)nOrb * nchunk;
        double* f = w + nOrb;
        double* df = f + nOrb;
        int* nit = reinterpret_cast<int*>(df + nOrb);
        int* done = nit + nOrb;
        int* active = done + nOrb;
        // qp_init writes w, f, nit and done before any read: nothing in the
        // shared array has to start at zero, whatever the last call left
        qp_init(s, eGWlin, w, f, nit, done, nOrb, nC, nR);
        for (int it = 0; it < QP_MAXIT; ++it) {
            sigma_partial(s, rhoT, false, w, eOld, Om, done, eta, doSRG, flow, nOrb, nC, nO, nR, nS, nchunk, part);
            qp_update(s, part, nchunk, eHF, w, f, df, nit, done, nOrb, QP_THRESH, QP_MAXIT, active);
        }
        qp_finish(s, w, df, nit, eGWlin, eOld, eGW, Z, nOrb, nC, nR, QP_MAXIT);
        return QK_OK;
    } catch (const Error& e) {
        g_err = e.msg;
        return e.code;
    } catch (...) {
        g_err = "unknown error";
        return QK_E_HIP;
    }
}

const char* qk_noalloc_last_error(void) { return qk::g_err.c_str(); }

}  // extern "C"
