// dump.hip -- field dumps written by a host thread while the run goes on.
//
// The reference writes every field at every dump (readwrite.f90:38-98:
// ./output/<name>_%08d.dat, the interior n^3 array in FP64, Fortran order,
// through MPI-IO) and the time loop waits for the files
// (main.f90:1011-1041).  mhit_dump_async keeps the file names and bytes
// and takes the writing off the time loop:
//   1. on the context's stream, in order with the timesteps, one kernel per
//      field copies it into a device staging buffer with the values of
//      mhit_download_field (u, v, w with the deferred mean removed, read
//      from the device scalars; p widened to FP64);
//   2. a host thread waits for that snapshot, copies each field to pinned
//      host memory on a stream of its own (two buffers: field f+1 moves
//      while field f is written) and writes the files.
// The call returns at once.  A new dump first waits for the previous one
// (the staging buffers are reused); mhit_dump_wait waits for the pending
// dump and returns its status.  Tracers and drop statistics are small and
// stay synchronous in the driver.
#include "internal.hpp"

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

struct mhit_dump_state {
    int device = 0;
    double* stage[MHIT_FIELD_RHSPHI_O + 1] = {};   // device snapshot per field id
    double* pinned[2] = {nullptr, nullptr};
    hipStream_t xfer = nullptr;
    hipEvent_t ready = nullptr;
    hipEvent_t moved[2] = {nullptr, nullptr};
    std::thread writer;
    bool running = false;
    int rc = MHIT_OK;
    std::string err;
};

using namespace mhit;

namespace {

const char* field_tag(int f) {
    static const char* tags[] = {"", "u", "v", "w", "p", "phi", "rhsu_o", "rhsv_o", "rhsw_o", "rhsphi_o"};
    return (f >= 1 && f <= MHIT_FIELD_RHSPHI_O) ? tags[f] : "";
}

// dst = src - S->mean[comp] (comp < 0: no mean), or the FP32 pressure widened
__global__ __launch_bounds__(256)
void k_snapshot(double* __restrict__ dst, const double* __restrict__ src, const float* __restrict__ srcf,
                const DevScalars* __restrict__ S, int comp, size_t n) {
    const double m = comp >= 0 ? S->mean[comp] : 0.0;
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        dst[q] = srcf ? (double)srcf[q] : (comp >= 0 ? src[q] - m : src[q]);
}

void write_files(mhit_dump_state* d, std::vector<int> fields, std::vector<std::string> paths, size_t N3) {
    auto fail = [&](int rc, const std::string& msg) {
        if (d->rc == MHIT_OK) { d->rc = rc; d->err = msg; }
    };
    if (hipSetDevice(d->device) != hipSuccess || hipEventSynchronize(d->ready) != hipSuccess) {
        fail(MHIT_ERR_HIP, "dump: waiting for the device snapshot failed");
        return;
    }
    const size_t bytes = N3 * sizeof(double);
    auto issue = [&](size_t i) -> bool {
        const int b = (int)(i & 1);
        return hipMemcpyAsync(d->pinned[b], d->stage[fields[i]], bytes, hipMemcpyDeviceToHost, d->xfer) == hipSuccess &&
               hipEventRecord(d->moved[b], d->xfer) == hipSuccess;
    };
    if (!issue(0)) { fail(MHIT_ERR_HIP, "dump: device-to-host copy failed"); return; }
    for (size_t i = 0; i < fields.size(); i++) {
        // the other buffer was written out in the previous iteration
        if (i + 1 < fields.size() && !issue(i + 1)) { fail(MHIT_ERR_HIP, "dump: device-to-host copy failed"); return; }
        const int b = (int)(i & 1);
        if (hipEventSynchronize(d->moved[b]) != hipSuccess) { fail(MHIT_ERR_HIP, "dump: copy wait failed"); return; }
        FILE* fp = std::fopen(paths[i].c_str(), "wb");
        if (!fp) { fail(MHIT_ERR_IO, "dump: cannot open " + paths[i] + ": " + std::strerror(errno)); continue; }
        const size_t w = std::fwrite(d->pinned[b], sizeof(double), N3, fp);
        const int cl = std::fclose(fp);
        if (w != N3 || cl != 0) fail(MHIT_ERR_IO, "dump: short write to " + paths[i]);
    }
}

int join(mhit_dump_state* d) {
    if (d->running) {
        d->writer.join();
        d->running = false;
    }
    const int rc = d->rc;
    if (rc) set_error(d->err);
    d->rc = MHIT_OK;
    d->err.clear();
    return rc;
}

}  // namespace

namespace mhit {

void dump_release(mhit_ctx* c) {
    mhit_dump_state* d = c->dump;
    if (!d) return;
    (void)join(d);
    for (double* s : d->stage)
        if (s) (void)hipFree(s);
    for (double* p : d->pinned)
        if (p) (void)hipHostFree(p);
    for (hipEvent_t e : d->moved)
        if (e) (void)hipEventDestroy(e);
    if (d->ready) (void)hipEventDestroy(d->ready);
    if (d->xfer) (void)hipStreamDestroy(d->xfer);
    delete d;
    c->dump = nullptr;
}

}  // namespace mhit

extern "C" {

int mhit_dump_async(mhit_ctx* c, const char* dir, int32_t step, uint32_t field_mask) {
    if (!c || !dir || step < 0) { set_error("bad arguments"); return MHIT_ERR_ARG; }
    std::vector<int> fields;
    for (int f = MHIT_FIELD_U; f <= MHIT_FIELD_RHSPHI_O; f++) {
        if (!(field_mask & (1u << f))) continue;
        const bool phase = f == MHIT_FIELD_PHI || f == MHIT_FIELD_RHSPHI_O;
        if (phase && !c->P.phiflag) { set_error("dump: phase-field fields need phiflag = 1"); return MHIT_ERR_ARG; }
        fields.push_back(f);
    }
    if (field_mask & ~(((1u << (MHIT_FIELD_RHSPHI_O + 1)) - 1) & ~1u)) {
        set_error("dump: unknown bits in the field mask");
        return MHIT_ERR_ARG;
    }
    if (!c->dump) {
        c->dump = new mhit_dump_state();
        c->dump->device = c->device;
    }
    mhit_dump_state* d = c->dump;
    int rc;
    if ((rc = join(d))) return rc;   // the previous dump (and its error, if any)
    if (fields.empty()) return MHIT_OK;
    const size_t N3 = c->N3;
    // transfer resources, created on first use; each one checked on its own
    // so that a call after a failed set-up completes it
    if (!d->xfer) MHIT_HIP_TRY(hipStreamCreateWithFlags(&d->xfer, hipStreamNonBlocking));
    if (!d->ready) MHIT_HIP_TRY(hipEventCreateWithFlags(&d->ready, hipEventDisableTiming));
    for (hipEvent_t& e : d->moved)
        if (!e) MHIT_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    for (double*& p : d->pinned)
        if (!p && hipHostMalloc((void**)&p, N3 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            set_error("dump: no pinned host memory for the transfer buffers");
            return MHIT_ERR_NOMEM;
        }
    std::vector<std::string> paths;
    for (int f : fields) {
        if (!d->stage[f] && hipMalloc((void**)&d->stage[f], N3 * sizeof(double)) != hipSuccess) {
            d->stage[f] = nullptr;
            set_error("dump: no device memory for the snapshot");
            return MHIT_ERR_NOMEM;
        }
        char name[64];
        std::snprintf(name, sizeof(name), "/%s_%08d.dat", field_tag(f), (int)step);
        paths.push_back(std::string(dir) + name);
        const double* src = nullptr;
        const float* srcf = nullptr;
        int comp = -1;
        if (f == MHIT_FIELD_P) {
            if (c->fp32) srcf = (const float*)c->pres;
            else src = (const double*)c->pres;
        } else {
            src = state_field(c, f);
            if (f <= MHIT_FIELD_W) comp = f - 1;
        }
        if (!src && !srcf) { set_error("dump: field not allocated"); return MHIT_ERR_ARG; }
        const int blocks = (int)std::min<size_t>((N3 + 255) / 256, 8192);
        k_snapshot<<<blocks, 256, 0, c->stream>>>(d->stage[f], src, srcf, c->S, comp, N3);
        MHIT_HIP_TRY(hipGetLastError());
    }
    MHIT_HIP_TRY(hipEventRecord(d->ready, c->stream));
    d->rc = MHIT_OK;
    d->writer = std::thread(write_files, d, fields, paths, N3);
    d->running = true;
    return MHIT_OK;
}

int mhit_dump_wait(mhit_ctx* c) {
    if (!c) { set_error("null argument"); return MHIT_ERR_ARG; }
    if (!c->dump) return MHIT_OK;
    return join(c->dump);
}

}  // extern "C"
