// Test infrastructure: binary record dump shared by the reference harness
// (ref_harness.cpp) and the CPU restatement (melody_oracle.cpp).
//
// A dump file is a sequence of records
//   u32 name_len | name bytes | u8 kind (0 = f64, 1 = i32) | u64 count | payload
// little-endian.  tests/mel_dump.py reads it into a dict of numpy arrays.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

struct DumpWriter {
    FILE* f = nullptr;
    explicit DumpWriter(const std::string& path) { f = std::fopen(path.c_str(), "wb"); }
    ~DumpWriter() { if (f) std::fclose(f); }
    bool ok() const { return f != nullptr; }
    void head(const std::string& name, uint8_t kind, uint64_t n) {
        uint32_t l = (uint32_t)name.size();
        std::fwrite(&l, 4, 1, f);
        std::fwrite(name.data(), 1, l, f);
        std::fwrite(&kind, 1, 1, f);
        std::fwrite(&n, 8, 1, f);
    }
    void f64(const std::string& name, const std::vector<double>& v) {
        head(name, 0, v.size());
        if (!v.empty()) std::fwrite(v.data(), 8, v.size(), f);
    }
    void i32(const std::string& name, const std::vector<int>& v) {
        head(name, 1, v.size());
        if (!v.empty()) std::fwrite(v.data(), 4, v.size(), f);
    }
    void scalar(const std::string& name, double x) { f64(name, std::vector<double>(1, x)); }
};
