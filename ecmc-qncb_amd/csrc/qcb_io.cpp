// qcb_io.cpp -- ILDG checkpoints staged through the device.
//
// save_ildg_clime / read_ildg_clime (reference src/io/ildg.cpp:34-237) in the
// file layout of host/lime.cpp, without a host-side copy of the field: the
// device converts its links to big-endian file-order doubles chunk by chunk
// (k_soa_to_be), the chunks stream over PCIe into two pinned host buffers, and
// the host writes one chunk while the device converts and copies the next.
// Loading runs the same pipeline backwards (k_be_to_soa).  The bytes written
// are those of qcb_ildg_write on the downloaded links.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "../host/lime.hpp"
#include "qcb_host.hpp"

using namespace qcb;

namespace {

constexpr uint64_t kLinkBytes = 144;          // 18 big-endian doubles
constexpr uint32_t kChunkLinks = 1u << 19;    // 75.5 MB per chunk

// Two device chunks, two pinned host chunks and their events; released on
// every return path.
struct Staging {
    qcb_ctx* ctx;
    uint64_t* dbuf[2] = {nullptr, nullptr};
    uint64_t* hbuf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    FILE* f = nullptr;
    explicit Staging(qcb_ctx* c) : ctx(c) {}
    int init(size_t bytes) {
        for (int b = 0; b < 2; ++b) {
            QCB_HIP(ctx, hipMalloc((void**)&dbuf[b], bytes));
            QCB_HIP(ctx, hipHostMalloc((void**)&hbuf[b], bytes, hipHostMallocDefault));
            QCB_HIP(ctx, hipEventCreateWithFlags(&ev[b], hipEventDisableTiming));
        }
        return QCB_OK;
    }
    ~Staging() {
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        for (int b = 0; b < 2; ++b) {
            if (dbuf[b]) (void)hipFree(dbuf[b]);
            if (hbuf[b]) (void)hipHostFree(hbuf[b]);
            if (ev[b]) (void)hipEventDestroy(ev[b]);
        }
        if (f) fclose(f);
    }
};

}  // namespace

extern "C" int qcb_ildg_save(qcb_lattice* lat, const char* path) {
    if (!lat || !path) return QCB_EINVAL;
    qcb_ctx* ctx = lat->ctx;
    const uint64_t n_links = (uint64_t)lat->V * 4;
    const uint32_t chunk = (uint32_t)std::min<uint64_t>(n_links, kChunkLinks);
    Staging st(ctx);
    int rc = st.init(chunk * kLinkBytes);
    if (rc) return rc;
    st.f = fopen(path, "wb");
    if (!st.f) return qcb_set_error(ctx, QCB_EINVAL, "cannot open %s for writing", path);
    const std::string xml = qcb_host::ildg_xml(lat->L);
    const auto h1 = qcb_host::lime_header("ildg-format", xml.size(), true, true);
    const auto h2 = qcb_host::lime_header("ildg-binary-data", n_links * kLinkBytes, false, true);
    static const unsigned char zeros[8] = {0};
    const size_t pad = (8 - xml.size() % 8) % 8;
    bool ok = fwrite(h1.data(), 1, 144, st.f) == 144 && fwrite(xml.data(), 1, xml.size(), st.f) == xml.size() &&
              fwrite(zeros, 1, pad, st.f) == pad && fwrite(h2.data(), 1, 144, st.f) == 144;
    if (!ok) return qcb_set_error(ctx, QCB_EINVAL, "short write to %s", path);
    const Launch g = lat_launch(lat, kFamLink);
    const uint64_t n_chunks = (n_links + chunk - 1) / chunk;
    uint32_t prev_n = 0;
    for (uint64_t c = 0; c <= n_chunks; ++c) {
        const int b = (int)(c & 1);
        uint32_t n = 0;
        if (c < n_chunks) {   // convert and copy chunk c
            const uint64_t lo = c * chunk;
            n = (uint32_t)std::min<uint64_t>(chunk, n_links - lo);
            if (lat->prec == QCB_F64) launch_soa_to_be<double>(g, (const double*)lat->U, st.dbuf[b], lo, n);
            else launch_soa_to_be<float>(g, (const float*)lat->U, st.dbuf[b], lo, n);
            QCB_HIP(ctx, hipGetLastError());
            QCB_HIP(ctx, hipMemcpyAsync(st.hbuf[b], st.dbuf[b], n * kLinkBytes, hipMemcpyDeviceToHost, ctx->stream));
            QCB_HIP(ctx, hipEventRecord(st.ev[b], ctx->stream));
        }
        if (c > 0) {          // write chunk c - 1 while chunk c is in flight
            QCB_HIP(ctx, hipEventSynchronize(st.ev[1 - b]));
            if (fwrite(st.hbuf[1 - b], 1, prev_n * kLinkBytes, st.f) != prev_n * kLinkBytes)
                return qcb_set_error(ctx, QCB_EINVAL, "short write to %s", path);
        }
        prev_n = n;
    }
    if (fclose(st.f) != 0) {
        st.f = nullptr;
        return qcb_set_error(ctx, QCB_EINVAL, "cannot close %s", path);
    }
    st.f = nullptr;
    return QCB_OK;
}

extern "C" int qcb_ildg_load(qcb_lattice* lat, const char* path) {
    if (!lat || !path) return QCB_EINVAL;
    qcb_ctx* ctx = lat->ctx;
    const uint64_t n_links = (uint64_t)lat->V * 4;
    std::vector<qcb_host::LimeRecord> recs;
    std::string err;
    if (!qcb_host::lime_scan(path, recs, err)) return qcb_set_error(ctx, QCB_EINVAL, "%s", err.c_str());
    const qcb_host::LimeRecord* bin = nullptr;
    for (const auto& r : recs)
        if (r.type == "ildg-binary-data") {
            bin = &r;
            break;
        }
    if (!bin) return qcb_set_error(ctx, QCB_EINVAL, "record 'ildg-binary-data' not found in %s", path);
    if (bin->length != n_links * kLinkBytes)
        return qcb_set_error(ctx, QCB_EINVAL, "ildg-binary-data holds %llu bytes, expected %llu",
                             (unsigned long long)bin->length, (unsigned long long)(n_links * kLinkBytes));
    const uint32_t chunk = (uint32_t)std::min<uint64_t>(n_links, kChunkLinks);
    Staging st(ctx);
    int rc = st.init(chunk * kLinkBytes);
    if (rc) return rc;
    st.f = fopen(path, "rb");
    if (!st.f || fseeko(st.f, (off_t)bin->offset, SEEK_SET) != 0)
        return qcb_set_error(ctx, QCB_EINVAL, "cannot seek in %s", path);
    const Launch g = lat_launch(lat, kFamLink);
    const uint64_t n_chunks = (n_links + chunk - 1) / chunk;
    for (uint64_t c = 0; c < n_chunks; ++c) {
        const int b = (int)(c & 1);
        const uint64_t lo = c * chunk;
        const uint32_t n = (uint32_t)std::min<uint64_t>(chunk, n_links - lo);
        // host buffer b was last read by the copy of chunk c - 2
        if (c >= 2) QCB_HIP(ctx, hipEventSynchronize(st.ev[b]));
        if (fread(st.hbuf[b], 1, n * kLinkBytes, st.f) != n * kLinkBytes)
            return qcb_set_error(ctx, QCB_EINVAL, "truncated binary data in %s", path);
        QCB_HIP(ctx, hipMemcpyAsync(st.dbuf[b], st.hbuf[b], n * kLinkBytes, hipMemcpyHostToDevice, ctx->stream));
        if (lat->prec == QCB_F64) launch_be_to_soa<double>(g, st.dbuf[b], (double*)lat->U, lo, n);
        else launch_be_to_soa<float>(g, st.dbuf[b], (float*)lat->U, lo, n);
        QCB_HIP(ctx, hipGetLastError());
        QCB_HIP(ctx, hipEventRecord(st.ev[b], ctx->stream));
    }
    QCB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QCB_OK;
}
