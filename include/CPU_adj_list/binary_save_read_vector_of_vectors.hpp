#pragma once
// Binary save/read of std::vector<std::vector<T>> -- the interface of
// GraphHET's binary_save_read_vector_of_vectors.hpp (:10-76), reimplemented.
//
// File layout (shared with ghet_host_graph_save / ghet_host_graph_load, which
// write a CSR as two such files): an int32 row count, then for every row an
// int32 length followed by that many T values, native byte order, no padding
// between records.  T must be trivially copyable.
//
// Errors follow the reference: a file that cannot be opened prints
// "Unable to open file <path>" plus a hint and exits with status 1; a short or
// malformed file prints "Unable to read file <path>" and exits with status 1.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <type_traits>
#include <vector>

namespace ghet_vv {

[[noreturn]] inline void fail_open(const std::string& path) {
    std::cout << "Unable to open file " << path << std::endl
              << "Please check the file location or file name." << std::endl;
    std::exit(1);
}

[[noreturn]] inline void fail_read(const std::string& path) {
    std::cout << "Unable to read file " << path << std::endl;
    std::exit(1);
}

struct File {
    std::FILE* f;
    explicit File(std::FILE* x) : f(x) {}
    ~File() {
        if (f) std::fclose(f);
    }
    File(const File&) = delete;
    File& operator=(const File&) = delete;
};

}  // namespace ghet_vv

template <typename T>
void binary_save_vector_of_vectors(const std::string& path, const std::vector<std::vector<T>>& myVector) {
    static_assert(std::is_trivially_copyable<T>::value, "rows must hold trivially copyable values");
    ghet_vv::File out(std::fopen(path.c_str(), "wb"));
    if (!out.f) ghet_vv::fail_open(path);
    const int32_t rows = static_cast<int32_t>(myVector.size());
    std::fwrite(&rows, sizeof(rows), 1, out.f);
    for (const std::vector<T>& r : myVector) {
        const int32_t len = static_cast<int32_t>(r.size());
        std::fwrite(&len, sizeof(len), 1, out.f);
        if (len > 0) std::fwrite(r.data(), sizeof(T), r.size(), out.f);
    }
}

template <typename T>
void binary_read_vector_of_vectors(const std::string& path, std::vector<std::vector<T>>& myVector) {
    static_assert(std::is_trivially_copyable<T>::value, "rows must hold trivially copyable values");
    myVector.clear();
    ghet_vv::File in(std::fopen(path.c_str(), "rb"));
    if (!in.f) ghet_vv::fail_open(path);
    int32_t rows = 0;
    if (std::fread(&rows, sizeof(rows), 1, in.f) != 1 || rows < 0) ghet_vv::fail_read(path);
    myVector.resize(static_cast<size_t>(rows));
    for (std::vector<T>& r : myVector) {
        int32_t len = 0;
        if (std::fread(&len, sizeof(len), 1, in.f) != 1 || len < 0) ghet_vv::fail_read(path);
        r.resize(static_cast<size_t>(len));
        if (len > 0 && std::fread(r.data(), sizeof(T), r.size(), in.f) != r.size()) ghet_vv::fail_read(path);
    }
}
