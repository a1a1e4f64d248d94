// Minimal .npz writer: a ZIP archive of uncompressed .npy (format 1.0)
// members, the container numpy.load reads and cnpy's npz_save writes (the
// reference saves its results this way, code/src/io/save.cpp:182-211).
#pragma once
#include <complex>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace cli {

class Npz {
  public:
    explicit Npz(const std::string &path) : f_(std::fopen(path.c_str(), "wb")) {}
    ~Npz() { close(); }
    bool ok() const { return f_ != nullptr; }

    void add(const std::string &name, const double *data, const std::vector<size_t> &shape)
    {
        member(name, "<f8", data, sizeof(double), shape);
    }
    void add(const std::string &name, const std::complex<double> *data, const std::vector<size_t> &shape)
    {
        member(name, "<c16", data, sizeof(std::complex<double>), shape);
    }
    void add(const std::string &name, const uint32_t *data, const std::vector<size_t> &shape)
    {
        member(name, "<u4", data, sizeof(uint32_t), shape);
    }

    bool close()
    {
        if (!f_) return false;
        const uint32_t cd_start = offset_;
        for (const auto &e : entries_) {
            put32(0x02014b50);
            put16(20);
            put16(20);
            put16(0);
            put16(0);
            put16(0);
            put16(0x21);  // 1980-01-01
            put32(e.crc);
            put32(e.size);
            put32(e.size);
            put16((uint16_t)e.name.size());
            put16(0);
            put16(0);
            put16(0);
            put16(0);
            put32(0);
            put32(e.offset);
            bytes(e.name.data(), e.name.size());
        }
        const uint32_t cd_size = offset_ - cd_start;
        put32(0x06054b50);
        put16(0);
        put16(0);
        put16((uint16_t)entries_.size());
        put16((uint16_t)entries_.size());
        put32(cd_size);
        put32(cd_start);
        put16(0);
        const bool good = std::fclose(f_) == 0 && !failed_;
        f_ = nullptr;
        return good;
    }

  private:
    struct Entry {
        std::string name;
        uint32_t crc, size, offset;
    };
    FILE *f_;
    uint32_t offset_ = 0;
    bool failed_ = false;
    std::vector<Entry> entries_;

    static uint32_t crc32(const std::vector<uint8_t> &b)
    {
        uint32_t c = 0xffffffffu;
        for (uint8_t x : b) {
            c ^= x;
            for (int k = 0; k < 8; k++) c = (c >> 1) ^ (0xedb88320u & (0u - (c & 1u)));
        }
        return ~c;
    }
    void bytes(const void *p, size_t n)
    {
        if (std::fwrite(p, 1, n, f_) != n) failed_ = true;
        offset_ += (uint32_t)n;
    }
    void put16(uint16_t v)
    {
        uint8_t b[2] = {(uint8_t)v, (uint8_t)(v >> 8)};
        bytes(b, 2);
    }
    void put32(uint32_t v)
    {
        uint8_t b[4] = {(uint8_t)v, (uint8_t)(v >> 8), (uint8_t)(v >> 16), (uint8_t)(v >> 24)};
        bytes(b, 4);
    }

    void member(const std::string &name, const char *descr, const void *data, size_t elem,
                const std::vector<size_t> &shape)
    {
        if (!f_) return;
        size_t count = 1;
        std::string sh = "(";
        for (size_t i = 0; i < shape.size(); i++) {
            count *= shape[i];
            sh += std::to_string(shape[i]) + (shape.size() == 1 ? "," : (i + 1 < shape.size() ? ", " : ""));
        }
        sh += ")";
        std::string hdr = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': " + sh + ", }";
        // magic (6) + version (2) + length (2) + header, padded with spaces to a multiple of 64, ending in '\n'
        size_t total = 10 + hdr.size() + 1;
        hdr += std::string((64 - total % 64) % 64, ' ') + "\n";
        std::vector<uint8_t> body;
        const char magic[] = "\x93NUMPY";
        body.insert(body.end(), magic, magic + 6);
        body.push_back(1);
        body.push_back(0);
        body.push_back((uint8_t)(hdr.size() & 0xff));
        body.push_back((uint8_t)(hdr.size() >> 8));
        body.insert(body.end(), hdr.begin(), hdr.end());
        const uint8_t *d = static_cast<const uint8_t *>(data);
        body.insert(body.end(), d, d + count * elem);
        const std::string fname = name + ".npy";
        Entry e{fname, crc32(body), (uint32_t)body.size(), offset_};
        put32(0x04034b50);
        put16(20);
        put16(0);
        put16(0);
        put16(0);
        put16(0x21);
        put32(e.crc);
        put32(e.size);
        put32(e.size);
        put16((uint16_t)fname.size());
        put16(0);
        bytes(fname.data(), fname.size());
        bytes(body.data(), body.size());
        entries_.push_back(e);
    }
};

}  // namespace cli
