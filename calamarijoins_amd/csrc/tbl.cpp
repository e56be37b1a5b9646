// tbl.cpp -- .tbl cache reader and writer (tbl.hpp).
#include "tbl.hpp"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <thread>

#include "cj_internal.hpp"

namespace cj {

namespace {

constexpr size_t kTblPage = 8192;

std::string errno_text(const std::string &what, const std::string &path) {
    return what + " " + path + ": " + std::strerror(errno);
}

void write_all(int fd, const void *p, size_t n, const std::string &path) {
    const char *c = static_cast<const char *>(p);
    while (n) {
        ssize_t w = ::write(fd, c, n);
        if (w < 0) {
            if (errno == EINTR) continue;
            throw InvalidError(errno_text("write failed:", path));
        }
        c += w;
        n -= (size_t)w;
    }
}

void pread_all(int fd, void *p, size_t n, uint64_t off) {
    char *c = static_cast<char *>(p);
    while (n) {
        ssize_t r = ::pread(fd, c, n, (off_t)off);
        if (r < 0) {
            if (errno == EINTR) continue;
            throw InvalidError(std::string("read failed: ") + std::strerror(errno));
        }
        if (r == 0) throw InvalidError("read past the end of the .tbl file");
        c += r;
        n -= (size_t)r;
        off += (uint64_t)r;
    }
}

}  // namespace

std::shared_ptr<TblFile> TblFile::open(const std::string &path, bool populate) {
    std::shared_ptr<TblFile> f(new TblFile());
    f->fd_ = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
    if (f->fd_ < 0) throw InvalidError(errno_text("cannot open", path));
    struct stat sb;
    if (fstat(f->fd_, &sb) != 0) throw InvalidError(errno_text("cannot stat", path));
    if (sb.st_size < (off_t)kTblPage) throw InvalidError("not a .tbl file (shorter than its header page): " + path);
    f->size_ = (size_t)sb.st_size;
    pread_all(f->fd_, &f->meta_, sizeof f->meta_, 0);
    const TblMeta &m = f->meta_;
    if (m.num_cols > kTblMaxColumns) throw InvalidError("a .tbl header holds at most 16 columns: " + path);
    uint64_t total = 0;
    for (uint64_t c = 0; c < m.num_cols; ++c) {
        if (m.types[c] < CJ_INT32 || m.types[c] > CJ_VARCHAR) throw InvalidError("unknown column type in " + path);
        if (m.num_pages[c] > (1ull << 32)) throw InvalidError("implausible page count in " + path);
        f->first_page_.push_back(total);
        total += m.num_pages[c];
    }
    if (f->size_ != kTblPage * (1 + total))
        throw InvalidError("size of " + path + " does not match its header (" + std::to_string(f->size_) +
                           " bytes for " + std::to_string(total) + " pages)");
    int flags = MAP_PRIVATE;
#ifdef MAP_POPULATE
    if (populate) flags |= MAP_POPULATE;
#endif
    f->map_ = mmap(nullptr, f->size_, PROT_READ, flags, f->fd_, 0);
    if (f->map_ == MAP_FAILED) {
        f->map_ = nullptr;
        throw InvalidError(errno_text("cannot map", path));
    }
    const uint8_t *data = static_cast<const uint8_t *>(f->map_) + kTblPage;
    f->pages_.resize(m.num_cols);
    for (uint64_t c = 0; c < m.num_cols; ++c) {
        auto &v = f->pages_[c];
        v.resize(m.num_pages[c]);
        for (uint64_t p = 0; p < m.num_pages[c]; ++p) v[p] = data + kTblPage * (f->first_page_[c] + p);
    }
    f->cols_.resize(m.num_cols);
    for (uint64_t c = 0; c < m.num_cols; ++c)
        f->cols_[c] = cj_column_desc{m.types[c], (uint32_t)m.num_pages[c], f->pages_[c].data()};
    f->desc_ = cj_table_desc{m.num_rows, (uint32_t)m.num_cols, f->cols_.data()};
    return f;
}

TblFile::~TblFile() {
    if (map_) munmap(map_, size_);
    if (fd_ >= 0) ::close(fd_);
}

void tbl_write(const std::string &path, const cj_table_desc &t) {
    if (t.num_columns > kTblMaxColumns) throw InvalidError("a .tbl file holds at most 16 columns");
    if (t.num_columns && !t.columns) throw InvalidError("null column array");
    TblMeta m{};
    m.num_rows = t.num_rows;
    m.num_cols = t.num_columns;
    for (uint32_t c = 0; c < t.num_columns; ++c) {
        const cj_column_desc &d = t.columns[c];
        if (d.type < CJ_INT32 || d.type > CJ_VARCHAR) throw InvalidError("unknown column type");
        if (d.num_pages && !d.pages) throw InvalidError("null page array");
        for (uint32_t p = 0; p < d.num_pages; ++p)
            if (!d.pages[p]) throw InvalidError("null page pointer");
        m.types[c] = d.type;
        m.num_pages[c] = d.num_pages;
    }
    const std::string tmp = path + ".tmp." + std::to_string(::getpid());
    int fd = ::open(tmp.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
    if (fd < 0) throw InvalidError(errno_text("cannot create", tmp));
    try {
        std::vector<uint8_t> buf(512 * kTblPage);  // 4 MiB writes
        std::memset(buf.data(), 0, kTblPage);
        std::memcpy(buf.data(), &m, sizeof m);
        size_t used = kTblPage;
        for (uint32_t c = 0; c < t.num_columns; ++c) {
            const cj_column_desc &d = t.columns[c];
            for (uint32_t p = 0; p < d.num_pages; ++p) {
                if (used == buf.size()) {
                    write_all(fd, buf.data(), used, tmp);
                    used = 0;
                }
                std::memcpy(buf.data() + used, d.pages[p], kTblPage);
                used += kTblPage;
            }
        }
        write_all(fd, buf.data(), used, tmp);
        if (::close(fd) != 0) {
            fd = -1;
            throw InvalidError(errno_text("close failed:", tmp));
        }
        fd = -1;
        if (std::rename(tmp.c_str(), path.c_str()) != 0) throw InvalidError(errno_text("cannot rename to", path));
    } catch (...) {
        if (fd >= 0) ::close(fd);
        ::unlink(tmp.c_str());
        throw;
    }
}

std::string tbl_cache_name(const char *table, const char *filter_text) {
    const char *f = filter_text ? filter_text : "No filter";
    uint64_t h = 14695981039346656037ull;  // FNV-1a 64 offset basis
    for (const unsigned char *p = reinterpret_cast<const unsigned char *>(f); *p; ++p) {
        h ^= *p;
        h *= 1099511628211ull;  // FNV 64 prime
    }
    return std::string(table ? table : "") + "_" + std::to_string(h) + ".tbl";
}

void tbl_read_pages(const TblFile &f, uint32_t c, uint32_t p0, uint32_t np, uint8_t *dst, int threads) {
    if (np == 0) return;
    threads = std::max(1, std::min<int>(threads, (int)((np + 255) / 256)));
    auto part = [&](uint32_t a, uint32_t b) {
        if (a < b) pread_all(f.fd(), dst + (size_t)a * kTblPage, (size_t)(b - a) * kTblPage, f.page_offset(c, p0 + a));
    };
    if (threads == 1) {
        part(0, np);
        return;
    }
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> err(threads);
    const uint32_t per = (np + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const uint32_t a = t * per, b = std::min(np, a + per);
        th.emplace_back([&, t, a, b] {
            try {
                part(a, b);
            } catch (...) {
                err[t] = std::current_exception();
            }
        });
    }
    for (auto &x : th) x.join();
    for (auto &e : err)
        if (e) std::rethrow_exception(e);
}

}  // namespace cj
