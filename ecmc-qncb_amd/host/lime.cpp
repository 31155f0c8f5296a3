// lime.cpp -- LIME / ILDG codec (see lime.hpp for the format).
#include "lime.hpp"

#include <cstdio>
#include <cstring>
#include <memory>

namespace qcb_host {

namespace {

constexpr uint32_t kLimeMagic = 0x456789abu;

void put_be(unsigned char* p, uint64_t v, int n) {
    for (int i = n - 1; i >= 0; --i) {
        p[i] = (unsigned char)(v & 0xff);
        v >>= 8;
    }
}

uint64_t get_be(const unsigned char* p, int n) {
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) v = (v << 8) | p[i];
    return v;
}

uint64_t pad8(uint64_t n) { return (8 - (n % 8)) % 8; }

struct FileCloser {
    void operator()(FILE* f) const {
        if (f) fclose(f);
    }
};
using File = std::unique_ptr<FILE, FileCloser>;

}  // namespace

std::array<unsigned char, 144> lime_header(const std::string& type, uint64_t length, bool mb, bool me) {
    std::array<unsigned char, 144> h{};
    put_be(h.data(), kLimeMagic, 4);
    put_be(h.data() + 4, 1, 2);
    put_be(h.data() + 6, (mb ? 0x8000u : 0u) | (me ? 0x4000u : 0u), 2);
    put_be(h.data() + 8, length, 8);
    std::memcpy(h.data() + 16, type.data(), type.size() < 127 ? type.size() : 127);
    return h;
}

bool lime_scan(const std::string& path, std::vector<LimeRecord>& out, std::string& err) {
    out.clear();
    File f(fopen(path.c_str(), "rb"));
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    uint64_t pos = 0;
    for (;;) {
        unsigned char h[144];
        size_t got = fread(h, 1, sizeof h, f.get());
        if (got == 0) return true;
        if (got != sizeof h || get_be(h, 4) != kLimeMagic) {
            err = "malformed LIME header at offset " + std::to_string(pos);
            return false;
        }
        LimeRecord r;
        uint64_t flags = get_be(h + 6, 2);
        r.mb = flags & 0x8000u;
        r.me = flags & 0x4000u;
        r.length = get_be(h + 8, 8);
        r.type.assign((const char*)h + 16, strnlen((const char*)h + 16, 128));
        r.offset = pos + 144;
        out.push_back(r);
        pos = r.offset + r.length + pad8(r.length);
        if (fseeko(f.get(), (off_t)pos, SEEK_SET) != 0) {
            err = "truncated record " + r.type;
            return false;
        }
    }
}

std::string ildg_xml(const int d[4]) {
    char buf[512];
    std::snprintf(buf, sizeof buf,
                  "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n"
                  "<ildgFormat><version>1.0</version><field>su3gauge</field>"
                  "<precision>64</precision><lx>%d</lx><ly>%d</ly><lz>%d</lz><lt>%d</lt>"
                  "</ildgFormat>",
                  d[0], d[1], d[2], d[3]);
    return std::string(buf);
}

bool ildg_write(const std::string& path, const int dims[4], const double* aos, std::string& err) {
    File f(fopen(path.c_str(), "wb"));
    if (!f) {
        err = "cannot open " + path + " for writing";
        return false;
    }
    const uint64_t V = (uint64_t)dims[0] * dims[1] * dims[2] * dims[3];
    const uint64_t n = V * 72;
    std::string xml = ildg_xml(dims);
    auto h1 = lime_header("ildg-format", xml.size(), true, true);
    static const unsigned char zeros[8] = {0};
    bool ok = fwrite(h1.data(), 1, 144, f.get()) == 144 && fwrite(xml.data(), 1, xml.size(), f.get()) == xml.size() &&
              fwrite(zeros, 1, pad8(xml.size()), f.get()) == pad8(xml.size());
    auto h2 = lime_header("ildg-binary-data", n * 8, false, true);
    ok = ok && fwrite(h2.data(), 1, 144, f.get()) == 144;
    std::vector<unsigned char> buf(1 << 20);
    for (uint64_t i = 0; ok && i < n;) {
        uint64_t m = std::min<uint64_t>(buf.size() / 8, n - i);
        for (uint64_t k = 0; k < m; ++k) {
            uint64_t bits;
            std::memcpy(&bits, aos + i + k, 8);
            put_be(buf.data() + 8 * k, bits, 8);
        }
        ok = fwrite(buf.data(), 1, m * 8, f.get()) == m * 8;
        i += m;
    }
    if (!ok) err = "short write to " + path;
    return ok;
}

bool ildg_read(const std::string& path, const int dims[4], double* aos, std::string& err) {
    std::vector<LimeRecord> recs;
    if (!lime_scan(path, recs, err)) return false;
    const uint64_t V = (uint64_t)dims[0] * dims[1] * dims[2] * dims[3];
    const uint64_t n = V * 72;
    for (const LimeRecord& r : recs) {
        if (r.type != "ildg-binary-data") continue;
        if (r.length != n * 8) {
            err = "ildg-binary-data holds " + std::to_string(r.length) + " bytes, expected " + std::to_string(n * 8);
            return false;
        }
        File f(fopen(path.c_str(), "rb"));
        if (!f || fseeko(f.get(), (off_t)r.offset, SEEK_SET) != 0) {
            err = "cannot seek in " + path;
            return false;
        }
        std::vector<unsigned char> buf(1 << 20);
        for (uint64_t i = 0; i < n;) {
            uint64_t m = std::min<uint64_t>(buf.size() / 8, n - i);
            if (fread(buf.data(), 1, m * 8, f.get()) != m * 8) {
                err = "truncated binary data in " + path;
                return false;
            }
            for (uint64_t k = 0; k < m; ++k) {
                uint64_t bits = get_be(buf.data() + 8 * k, 8);
                std::memcpy(aos + i + k, &bits, 8);
            }
            i += m;
        }
        return true;
    }
    err = "record 'ildg-binary-data' not found in " + path;
    return false;
}

}  // namespace qcb_host
