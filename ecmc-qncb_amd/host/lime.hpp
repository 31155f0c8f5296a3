// lime.hpp -- LIME container records and the ILDG gauge-configuration format,
// written from the public LIME / ILDG specifications (no c-lime dependency).
//
// LIME record = 144-byte header + payload padded with zeros to a multiple of 8:
//   magic 0x456789ab (u32 BE) | version 1 (u16 BE) | flags (u16 BE: bit15 = MB,
//   bit14 = ME) | payload length (u64 BE) | type (128 bytes, NUL padded).
// ILDG file as the reference writes it (src/io/ildg.cpp:13-22,34-136):
//   record "ildg-format" (MB=1, ME=1) with the XML below, then record
//   "ildg-binary-data" (MB=0, ME=1) holding V*4*9 complex big-endian doubles
//   in t, z, y, x (x fastest), mu, row-major 3x3 order.
// The reader accepts any MB/ME combination and skips unknown records.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace qcb_host {

struct LimeRecord {
    std::string type;
    uint64_t length = 0;   // payload bytes (before padding)
    uint64_t offset = 0;   // file offset of the payload
    bool mb = false, me = false;
};

// Header bytes of one record (exposed for tests of the layout).
std::array<unsigned char, 144> lime_header(const std::string& type, uint64_t length, bool mb, bool me);

// Lists every record of a LIME file; returns false on a malformed header.
bool lime_scan(const std::string& path, std::vector<LimeRecord>& out, std::string& err);

// The ildg-format XML of the reference, generalised to Lx, Ly, Lz, Lt.
std::string ildg_xml(const int dims[4]);

// Writes an ILDG file from host-endian AoS doubles (V*72 values).
bool ildg_write(const std::string& path, const int dims[4], const double* aos, std::string& err);
// Reads the ildg-binary-data record (size checked against dims) into aos.
bool ildg_read(const std::string& path, const int dims[4], double* aos, std::string& err);

}  // namespace qcb_host
