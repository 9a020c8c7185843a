// risgraph/io.hpp -- binary edge files, as read by the reference drivers
// (core/io.hpp:26-36): a file of (u64 src, u64 dst) pairs in host byte order.
#pragma once
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

static inline std::pair<std::pair<uint64_t, uint64_t> *, uint64_t> mmap_binary(const std::string &path)
{
    int fd = open(path.c_str(), O_RDONLY);
    if (fd < 0) throw std::runtime_error("open path " + path + " error.");
    struct stat st;
    if (fstat(fd, &st) != 0) {
        close(fd);
        throw std::runtime_error("stat " + path + " error.");
    }
    size_t size = (size_t)st.st_size;
    void *p = size ? mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0) : nullptr;
    close(fd);
    if (size && p == MAP_FAILED) throw std::runtime_error("mmap error.");
    return {(std::pair<uint64_t, uint64_t> *)p, size / sizeof(std::pair<uint64_t, uint64_t>)};
}
