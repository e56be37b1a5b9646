// tbl.hpp -- the harness's .tbl table cache: one 8192-byte header page
// (TableMeta: num_rows, num_cols, types[16], num_pages[16], zero filler)
// followed by every column's 8 KiB pages, column after column (reference
// include/data_access/table.h:30-37 and DumpTable::dump, table.h:208-221;
// reader Table::from_cache, src/build_table.cpp:140-197).
#pragma once
#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cj_api.h"

namespace cj {

// Header layout: DataType is a 4-byte enum, so num_pages starts at byte 80.
struct TblMeta {
    uint64_t num_rows;
    uint64_t num_cols;
    int32_t types[16];
    uint64_t num_pages[16];
};
static_assert(sizeof(TblMeta) == 208, "TableMeta layout");

constexpr uint32_t kTblMaxColumns = 16;

// A read-only mapping of a .tbl file with the descriptor arrays that point
// into it.
class TblFile {
  public:
    static std::shared_ptr<TblFile> open(const std::string &path, bool populate);
    ~TblFile();
    const cj_table_desc &desc() const { return desc_; }
    const TblMeta &meta() const { return meta_; }
    int fd() const { return fd_; }
    // byte offset of page p of column c in the file
    uint64_t page_offset(uint32_t c, uint32_t p) const { return 8192ull * (1 + first_page_[c] + p); }

  private:
    TblFile() = default;
    int fd_ = -1;
    void *map_ = nullptr;
    size_t size_ = 0;
    TblMeta meta_{};
    std::vector<uint64_t> first_page_;
    std::vector<std::vector<const void *>> pages_;
    std::vector<cj_column_desc> cols_;
    cj_table_desc desc_{};
};

// DumpTable::dump: header page, then the pages of each column in order.
// Written to a temporary file that is renamed over `path`.
void tbl_write(const std::string &path, const cj_table_desc &t);

// The harness's cache file name for a scan: "<table>_<FNV-1a 64 of the
// filter text>.tbl" (tests/read_sql.cpp:1156-1158, table.h:15-27).
std::string tbl_cache_name(const char *table, const char *filter_text);

// Reads pages [p0, p0 + np) of column c into dst with parallel preads.
void tbl_read_pages(const TblFile &f, uint32_t c, uint32_t p0, uint32_t np, uint8_t *dst, int threads);

}  // namespace cj
