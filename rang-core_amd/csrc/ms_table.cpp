// casacore table reader (host side of Measurement Set ingest). The on-disk
// layouts are the ones core/src/ms.rs reads (the file:line citations name
// where); the code is this library's own.
#include "ms_table.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <unordered_map>

namespace rang {
namespace ms {
namespace {

constexpr std::uint32_t kMagic = 0xbebebebeu;
constexpr std::uint64_t kBucketStart = 512;  // buckets follow a 512-byte header
constexpr std::size_t kTileCache = 64;

[[noreturn]] void fail(const std::string& m) { throw Error(m); }

std::string fmt_u(std::uint64_t v) { return std::to_string(v); }

std::vector<std::uint8_t> slurp(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail("io: cannot open " + path);
    return std::vector<std::uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

// A file opened once and read at offsets.
class File {
public:
    explicit File(const std::string& p) : path_(p), f_(std::fopen(p.c_str(), "rb")) {
        if (!f_) fail("io: cannot open " + p);
    }
    ~File() {
        if (f_) std::fclose(f_);
    }
    File(const File&) = delete;
    File& operator=(const File&) = delete;
    std::vector<std::uint8_t> at(std::uint64_t off, std::size_t n) {
        std::vector<std::uint8_t> out(n);
        if (n == 0) return out;
        if (std::fseek(f_, long(off), SEEK_SET) != 0 || std::fread(out.data(), 1, n, f_) != n)
            fail("read of " + fmt_u(n) + " bytes at " + fmt_u(off) + " of " + path_ + " failed");
        return out;
    }
    std::uint64_t size() {
        std::fseek(f_, 0, SEEK_END);
        return std::uint64_t(std::ftell(f_));
    }

private:
    std::string path_;
    std::FILE* f_;
};

std::uint32_t get32(const std::uint8_t* p, bool big) {
    return big ? (std::uint32_t(p[0]) << 24 | std::uint32_t(p[1]) << 16 | std::uint32_t(p[2]) << 8 | p[3])
               : (std::uint32_t(p[3]) << 24 | std::uint32_t(p[2]) << 16 | std::uint32_t(p[1]) << 8 | p[0]);
}
std::uint64_t get64(const std::uint8_t* p, bool big) {
    const std::uint64_t a = get32(p, big), b = get32(p + 4, big);
    return big ? (a << 32 | b) : (b << 32 | a);
}

// Bounds-checked reads of a bucket.
std::uint32_t u32_at(const std::vector<std::uint8_t>& b, std::size_t at, bool big) {
    if (at + 4 > b.size()) fail("truncated bucket");
    return get32(b.data() + at, big);
}
std::uint64_t u64_at(const std::vector<std::uint8_t>& b, std::size_t at, bool big) {
    if (at + 8 > b.size()) fail("truncated bucket");
    return get64(b.data() + at, big);
}

// An AipsIO reading position (the serialisation of table.dat and the
// managers' headers; core/src/ms.rs:200-340): objects are a length, a type
// name and a version; integers canonical big-endian or little-endian.
struct Aips {
    const std::uint8_t* p;
    std::size_t n, pos = 0;
    bool big;
    Aips(const std::vector<std::uint8_t>& b, bool be) : p(b.data()), n(b.size()), big(be) {}
    const std::uint8_t* take(std::size_t k) {
        if (pos + k > n) fail("truncated AipsIO object at byte " + fmt_u(pos));
        const std::uint8_t* q = p + pos;
        pos += k;
        return q;
    }
    std::uint8_t u8() { return *take(1); }
    bool boolean() { return u8() != 0; }
    std::uint32_t u32() { return get32(take(4), big); }
    std::int32_t i32() { return std::int32_t(u32()); }
    std::uint64_t u64() { return get64(take(8), big); }
    std::string str() {
        const std::uint32_t k = u32();
        const std::uint8_t* q = take(k);
        return std::string(reinterpret_cast<const char*>(q), k);
    }
    void magic() {
        if (u32() != kMagic) fail("AipsIO magic value missing");
    }
    struct Obj {
        std::size_t start, length;
        std::string kind;
        std::uint32_t version;
    };
    Obj begin() {
        Obj o;
        o.start = pos;
        o.length = u32();
        o.kind = str();
        o.version = u32();
        return o;
    }
    Obj begin(const char* kind) {
        Obj o = begin();
        if (o.kind != kind) fail(std::string("expected AipsIO object ") + kind + ", found " + o.kind);
        return o;
    }
    void end(const Obj& o) { pos = o.start + o.length; }
    void skip() { end(begin()); }
    std::vector<std::int64_t> ipos() {
        Obj o = begin("IPosition");
        const std::uint32_t k = u32();
        std::vector<std::int64_t> v(k);
        for (auto& x : v) {
            if (o.version == 1)
                x = i32();
            else if (o.version == 2)
                x = std::int64_t(u64());
            else
                fail("IPosition version " + fmt_u(o.version));
        }
        end(o);
        return v;
    }
    std::vector<std::uint64_t> block(bool wide) {
        Obj o = begin("Block");
        const std::uint32_t k = u32();
        std::vector<std::uint64_t> v(k);
        for (auto& x : v) x = wide ? u64() : u32();
        end(o);
        return v;
    }
};

std::size_t product(const std::vector<std::int64_t>& s) {
    std::size_t k = 1;
    for (auto x : s) k *= std::size_t(std::max<std::int64_t>(x, 0));
    return k;
}

// `count` values of `d` from file bytes into native order.
void append_values(std::vector<std::uint8_t>& out, const std::uint8_t* src, std::size_t avail, DType d,
                   std::size_t count, bool big) {
    const std::size_t sz = dtype_size(d), nb = sz * count;
    if (nb > avail) fail("truncated value run");
    const std::size_t at = out.size();
    out.insert(out.end(), src, src + nb);
    if (big && sz > 1) {
        // complex values swap as two reals
        const std::size_t unit = d == kComplex ? 4 : d == kDComplex ? 8 : sz;
        for (std::size_t i = at; i < out.size(); i += unit) std::reverse(out.begin() + i, out.begin() + i + unit);
    }
}

// `count` booleans stored as bits from bit `bit0` (least significant first).
void append_bits(std::vector<std::uint8_t>& out, const std::uint8_t* src, std::size_t avail, std::size_t bit0,
                 std::size_t count) {
    if (count == 0) return;
    if (bit0 + count > avail * 8) fail("truncated boolean bits");
    for (std::size_t i = 0; i < count; ++i) {
        const std::size_t b = bit0 + i;
        out.push_back((src[b >> 3] >> (b & 7)) & 1);
    }
}

std::string text(const std::uint8_t* p, std::size_t n) { return std::string(reinterpret_cast<const char*>(p), n); }

struct TableFiles {
    std::string dir;
    bool little;
    std::uint64_t nrows;
    std::string file(std::uint32_t seq, const std::string& suffix = "") const {
        return dir + "/table.f" + std::to_string(seq) + suffix;
    }
};

// Runs of consecutive rows of a row list, in list order.
template <class F>
void for_runs(const std::vector<std::uint64_t>& rows, F&& f) {
    std::size_t i = 0;
    while (i < rows.size()) {
        std::size_t k = 1;
        while (i + k < rows.size() && rows[i + k] == rows[i] + k) ++k;
        f(rows[i], std::uint64_t(k));
        i += k;
    }
}

// Cells read whole, cut to a region (core/src/ms.rs:1728-1816).
Cells cut_cells(Cells in, const Slice& cut) {
    const Slice reg = resolve_slice(cut, in.shape);
    const int nd = int(in.shape.size());
    std::vector<std::int64_t> len(nd);
    for (int i = 0; i < nd; ++i) len[i] = reg[i].second;
    Cells out;
    out.dtype = in.dtype;
    out.nrows = in.nrows;
    out.shape = len;
    const std::size_t cell = product(in.shape), ocell = product(len);
    const bool str = in.dtype == kString;
    const std::size_t sz = str ? 0 : native_size(in.dtype);
    std::vector<std::size_t> stride(nd, 1);
    for (int i = 1; i < nd; ++i) stride[i] = stride[i - 1] * std::size_t(in.shape[i - 1]);
    for (std::uint64_t r = 0; r < in.nrows; ++r) {
        for (std::size_t e = 0; e < ocell; ++e) {
            std::size_t rem = e, at = r * cell;
            for (int i = 0; i < nd; ++i) {
                const std::size_t k = rem % std::size_t(len[i]);
                rem /= std::size_t(len[i]);
                at += (std::size_t(reg[i].first) + k) * stride[i];
            }
            if (str)
                out.strings.push_back(in.strings[at]);
            else
                out.bytes.insert(out.bytes.end(), in.bytes.begin() + at * sz, in.bytes.begin() + (at + 1) * sz);
        }
    }
    return out;
}

// Appends `part` (rows of one shape) to `acc`; the first part sets the shape.
void merge(Cells& acc, bool& first, Cells&& part, const std::string& name) {
    if (first) {
        acc = std::move(part);
        first = false;
        return;
    }
    if (acc.shape != part.shape) fail("rows of " + name + " differ in shape; read fewer rows at once");
    acc.nrows += part.nrows;
    acc.bytes.insert(acc.bytes.end(), part.bytes.begin(), part.bytes.end());
    for (auto& s : part.strings) acc.strings.push_back(std::move(s));
}

// ------------------------------------------------------------ indirect arrays

// table.fNi, the arrays of non-direct columns (StManArrayFile;
// core/src/ms.rs:966-1059): a version word, then per array an optional
// reference count, the dimension count, the dimensions and the values.
class ArrayFile {
public:
    ArrayFile(const std::string& data_file, bool big) : f_(data_file + "i"), big_(big) {
        version_ = get32(f_.at(0, 4).data(), big_);
    }
    std::vector<std::int64_t> shape(std::uint64_t off) {
        std::uint64_t at = off + (version_ > 0 ? 4 : 0);
        const std::uint32_t nd = get32(f_.at(at, 4).data(), big_);
        if (nd > 64) fail("indirect array with " + fmt_u(nd) + " dimensions");
        at += 4;
        const auto dims = f_.at(at, 4 * std::size_t(nd));
        std::vector<std::int64_t> s(nd);
        for (std::uint32_t i = 0; i < nd; ++i) s[i] = std::int32_t(get32(dims.data() + 4 * i, big_));
        data_at_ = at + 4 * std::uint64_t(nd);
        return s;
    }
    // The values of the array whose shape was read last.
    void values(DType d, const std::vector<std::int64_t>& s, Cells& out) {
        const std::size_t n = product(s);
        if (d == kBool) {
            const auto b = f_.at(data_at_, (n + 7) / 8);
            append_bits(out.bytes, b.data(), b.size(), 0, n);
        } else if (d == kString) {
            const auto offs = f_.at(data_at_, 4 * n);
            for (std::size_t i = 0; i < n; ++i) {
                const std::uint64_t o = get32(offs.data() + 4 * i, big_);
                if (o == 0) {
                    out.strings.emplace_back();
                    continue;
                }
                const std::uint32_t len = get32(f_.at(o, 4).data(), big_);
                const auto t = f_.at(o + 4, len);
                out.strings.push_back(text(t.data(), t.size()));
            }
        } else {
            const auto b = f_.at(data_at_, dtype_size(d) * n);
            append_values(out.bytes, b.data(), b.size(), d, n, big_);
        }
    }

private:
    File f_;
    bool big_;
    std::uint32_t version_ = 0;
    std::uint64_t data_at_ = 0;
};

// One indirect cell per row offset into the array file.
void read_indirect(const std::string& data_file, bool big, const ColumnDesc& d,
                   const std::vector<std::uint64_t>& offsets, std::uint64_t first_row, Cells& out,
                   std::unique_ptr<ArrayFile>& af, bool& have_shape) {
    for (std::size_t i = 0; i < offsets.size(); ++i) {
        if (offsets[i] == 0) fail("row " + fmt_u(first_row + i) + " of " + d.name + " has no array");
        if (!af) af.reset(new ArrayFile(data_file, big));
        const auto s = af->shape(offsets[i]);
        if (!have_shape) {
            out.shape = s;
            have_shape = true;
        } else if (s != out.shape) {
            fail("rows of " + d.name + " differ in shape; read fewer rows at once");
        }
        af->values(d.dtype, s, out);
    }
}

// ------------------------------------------------------------ StandardStMan

// StandardStMan (core/src/ms.rs:434-964): a header in the first 512 bytes,
// fixed-size buckets after it, an index (last row, bucket) per column group
// in a chain of buckets, each column's per-row slots at an offset within a
// bucket, variable-length strings in string buckets.
class Ssm : public Manager {
public:
    Ssm(const TableFiles& t, std::uint32_t seq, const std::vector<std::uint8_t>& blob,
        const std::vector<const ColumnDesc*>& cols)
        : file_(t.file(seq)) {
        Aips c(blob, true);
        c.magic();
        auto o = c.begin("SSM");
        (void)c.str();
        col_offset_ = c.block(false);
        col_index_ = c.block(false);
        c.end(o);
        File f(file_);
        const auto head = f.at(0, kBucketStart);
        Aips h(head, !t.little);
        h.magic();
        auto ho = h.begin("StandardStMan");
        big_ = ho.version >= 3 ? h.boolean() : true;
        if (big_ == t.little) fail("StandardStMan endian flag mismatches the table");
        bucket_ = h.u32();
        nbuckets_ = h.u32();
        (void)h.u32();  // cache size
        (void)h.u32();  // free buckets
        (void)h.i32();  // first free
        const std::uint32_t nidx_buckets = h.u32();
        std::int32_t ib = h.i32();
        const std::uint32_t idx_off = ho.version >= 2 ? h.u32() : 0;
        (void)h.i32();  // last string bucket
        std::size_t left = h.u32();
        const std::uint32_t nindex = h.u32();
        // the index bytes: after two words per chain bucket, or at idx_off
        std::vector<std::uint8_t> ix;
        for (std::uint32_t k = 0; k < nidx_buckets; ++k) {
            if (ib < 0) fail("StandardStMan index bucket chain ends early");
            const auto b = f.at(kBucketStart + std::uint64_t(ib) * bucket_, bucket_);
            const std::int32_t next = std::int32_t(u32_at(b, 4, true));
            std::size_t from = 8, take;
            if (idx_off > 0) {
                from = idx_off;
                take = left;
            } else {
                take = std::min(left, std::size_t(bucket_) - 8);
            }
            if (from + take > b.size()) fail("StandardStMan index overruns its bucket");
            ix.insert(ix.end(), b.begin() + from, b.begin() + from + take);
            left = left > bucket_ - 8 ? left - (bucket_ - 8) : 0;
            ib = next;
        }
        Aips ic(ix, big_);
        for (std::uint32_t k = 0; k < nindex; ++k) {
            ic.magic();
            auto io = ic.begin("SSMIndex");
            const std::uint32_t used = ic.u32();
            (void)ic.u32();  // rows per bucket
            (void)ic.u32();  // columns
            ic.skip();       // free-space map
            Index x;
            x.last = ic.block(io.version != 1);
            auto bk = ic.block(false);
            x.last.resize(std::min<std::size_t>(used, x.last.size()));
            bk.resize(std::min<std::size_t>(used, bk.size()));
            x.bucket.assign(bk.begin(), bk.end());
            index_.push_back(std::move(x));
            ic.end(io);
        }
        for (const ColumnDesc* d : cols) {
            Col k;
            k.dtype = d->dtype;
            const std::size_t ne = product(d->shape);
            if (!d->array) {
                k.kind = d->dtype == kString ? (d->max_length > 0 ? kFixedStr : kIndStr) : kScalar;
                k.width = d->max_length;
            } else if (d->direct()) {
                k.kind = kDirect;
                k.nelem = ne;
            } else {
                k.kind = d->dtype == kString ? kIndStrArray : kIndirect;
            }
            cols_.push_back(k);
        }
        if (col_offset_.size() != cols_.size() || col_index_.size() != cols_.size())
            fail("StandardStMan column count mismatches the table");
    }

    Cells read(const ColumnDesc& d, int col, const std::vector<std::uint64_t>& rows, const Slice* cut) override {
        Cells acc;
        acc.dtype = d.dtype;
        bool first = true;
        for_runs(rows, [&](std::uint64_t s, std::uint64_t n) { merge(acc, first, run(d, col, s, n), d.name); });
        if (first) acc.nrows = 0;
        if (cut && acc.nrows > 0) return cut_cells(std::move(acc), *cut);
        return acc;
    }

private:
    enum Kind { kScalar, kFixedStr, kIndStr, kDirect, kIndirect, kIndStrArray };
    struct Col {
        DType dtype;
        Kind kind;
        std::size_t nelem = 0;
        std::uint32_t width = 0;
    };
    struct Index {
        std::vector<std::uint64_t> last;
        std::vector<std::uint32_t> bucket;
    };
    std::string file_;
    bool big_ = true;
    std::uint32_t bucket_ = 0, nbuckets_ = 0;
    std::vector<std::uint64_t> col_offset_, col_index_;
    std::vector<Index> index_;
    std::vector<Col> cols_;

    std::vector<std::uint8_t> bucket(File& f, std::uint32_t b) const {
        if (b >= nbuckets_) fail("StandardStMan bucket " + fmt_u(b) + " beyond the file");
        return f.at(kBucketStart + std::uint64_t(b) * bucket_, bucket_);
    }

    // Visits rows s .. s + n of `col` as (bucket bytes, slot offset, row in
    // bucket, rows) runs.
    template <class F>
    void visit(File& f, int col, std::uint64_t s, std::uint64_t n, F&& fn) const {
        const Index& x = index_[col_index_[col]];
        std::uint64_t row = s;
        while (row < s + n) {
            const std::size_t i = std::size_t(std::lower_bound(x.last.begin(), x.last.end(), row) - x.last.begin());
            if (i >= x.last.size()) fail("row " + fmt_u(row) + " beyond the StandardStMan index");
            const std::uint64_t first = i == 0 ? 0 : x.last[i - 1] + 1;
            const std::uint64_t last = std::min(x.last[i], s + n - 1);
            const auto b = bucket(f, x.bucket[i]);
            fn(b, std::size_t(col_offset_[col]), std::size_t(row - first), std::size_t(last - row + 1));
            row = last + 1;
        }
    }

    // String buckets: 16 bytes of canonical bookkeeping (bucket, used
    // length, deleted, next bucket), then characters; a string continues
    // in the next bucket of the chain (core/src/ms.rs:850-964).
    struct Strings {
        const Ssm& m;
        File f;
        std::int64_t cur = -1;
        std::vector<std::uint8_t> b;
        std::size_t off = 0;
        explicit Strings(const Ssm& s) : m(s), f(s.file_) {}
        void seek(std::int32_t bk, std::size_t o) {
            if (bk != cur) {
                if (bk < 0) fail("string bucket chain ends early");
                b = m.bucket(f, std::uint32_t(bk));
                cur = bk;
            }
            off = o;
        }
        std::string read(std::size_t len) {
            std::string out;
            std::uint32_t guard = 0;
            while (out.size() < len) {
                const std::int32_t used_i = std::int32_t(u32_at(b, 4, true));
                const std::size_t used = used_i > 0 ? std::size_t(used_i) : 0;
                const std::size_t take = std::min(len - out.size(), used > off ? used - off : 0);
                if (16 + off + take > b.size()) fail("string bucket overrun");
                out.append(reinterpret_cast<const char*>(b.data()) + 16 + off, take);
                off += take;
                if (out.size() < len) {
                    if (++guard > m.nbuckets_) fail("string bucket chain loops");
                    seek(std::int32_t(u32_at(b, 12, true)), 0);
                }
            }
            return out;
        }
        std::int32_t integer() {
            const std::string s = read(4);
            return std::int32_t(get32(reinterpret_cast<const std::uint8_t*>(s.data()), true));
        }
        void strings(std::size_t n, std::vector<std::string>& out) {
            for (std::size_t i = 0; i < n; ++i) {
                const std::int32_t len = integer();
                out.push_back(read(std::size_t(std::max(len, 0))));
            }
        }
    };

    Cells run(const ColumnDesc& d, int col, std::uint64_t s, std::uint64_t n) const {
        File f(file_);
        const Col& c = cols_[col];
        const bool big = big_;
        Cells out;
        out.dtype = d.dtype;
        out.nrows = n;
        switch (c.kind) {
        case kScalar:
            visit(f, col, s, n, [&](const std::vector<std::uint8_t>& b, std::size_t o, std::size_t r, std::size_t k) {
                if (o > b.size()) fail("slot outside its bucket");
                if (c.dtype == kBool)
                    append_bits(out.bytes, b.data() + o, b.size() - o, r, k);
                else {
                    const std::size_t at = o + r * dtype_size(c.dtype);
                    if (at > b.size()) fail("slot outside its bucket");
                    append_values(out.bytes, b.data() + at, b.size() - at, c.dtype, k, big);
                }
            });
            break;
        case kFixedStr:
            visit(f, col, s, n, [&](const std::vector<std::uint8_t>& b, std::size_t o, std::size_t r, std::size_t k) {
                for (std::size_t i = 0; i < k; ++i) {
                    const std::size_t at = o + (r + i) * c.width;
                    if (at + c.width > b.size()) fail("slot outside its bucket");
                    std::size_t len = 0;
                    while (len < c.width && b[at + len] != 0) ++len;
                    out.strings.push_back(text(b.data() + at, len));
                }
            });
            break;
        case kIndStr: {
            Strings h(*this);
            visit(f, col, s, n, [&](const std::vector<std::uint8_t>& b, std::size_t o, std::size_t r, std::size_t k) {
                for (std::size_t i = 0; i < k; ++i) {
                    const std::size_t at = o + (r + i) * 12;
                    const std::int32_t bk = std::int32_t(u32_at(b, at, big)), of = std::int32_t(u32_at(b, at + 4, big)),
                                       ln = std::int32_t(u32_at(b, at + 8, big));
                    if (ln <= 8) {
                        // short strings sit in the slot itself
                        out.strings.push_back(text(b.data() + at, std::size_t(std::max(ln, 0))));
                    } else {
                        h.seek(bk, std::size_t(of));
                        out.strings.push_back(h.read(std::size_t(ln)));
                    }
                }
            });
            break;
        }
        case kDirect: {
            out.shape = d.shape;
            Strings h(*this);
            visit(f, col, s, n, [&](const std::vector<std::uint8_t>& b, std::size_t o, std::size_t r, std::size_t k) {
                if (c.dtype == kBool) {
                    if (o > b.size()) fail("slot outside its bucket");
                    append_bits(out.bytes, b.data() + o, b.size() - o, r * c.nelem, k * c.nelem);
                } else if (c.dtype == kString) {
                    for (std::size_t i = 0; i < k; ++i) {
                        const std::size_t at = o + (r + i) * 12;
                        const std::int32_t bk = std::int32_t(u32_at(b, at, big)),
                                           of = std::int32_t(u32_at(b, at + 4, big)),
                                           ln = std::int32_t(u32_at(b, at + 8, big));
                        if (ln <= 0) {
                            out.strings.resize(out.strings.size() + c.nelem);
                        } else {
                            h.seek(bk, std::size_t(of));
                            h.strings(c.nelem, out.strings);
                        }
                    }
                } else {
                    const std::size_t at = o + r * c.nelem * dtype_size(c.dtype);
                    if (at > b.size()) fail("slot outside its bucket");
                    append_values(out.bytes, b.data() + at, b.size() - at, c.dtype, k * c.nelem, big);
                }
            });
            break;
        }
        case kIndStrArray: {
            // a (bucket, offset, length) slot per row into the string
            // buckets; a variable shape stores its shape and a filled flag
            // first (core/src/ms.rs:453-462, 718-781)
            Strings h(*this);
            bool have = false;
            std::uint64_t row = s;
            visit(f, col, s, n, [&](const std::vector<std::uint8_t>& b, std::size_t o, std::size_t r, std::size_t k) {
                for (std::size_t i = 0; i < k; ++i, ++row) {
                    const std::size_t at = o + (r + i) * 12;
                    const std::int32_t bk = std::int32_t(u32_at(b, at, big)), of = std::int32_t(u32_at(b, at + 4, big)),
                                       ln = std::int32_t(u32_at(b, at + 8, big));
                    std::vector<std::int64_t> shape;
                    bool filled;
                    if (d.fixed) {
                        shape = d.shape;
                        filled = ln > 0;
                        if (filled) h.seek(bk, std::size_t(of));
                    } else {
                        if (ln <= 0) fail("row " + fmt_u(row) + " of " + d.name + " has no array");
                        h.seek(bk, std::size_t(of));
                        const std::int32_t nd = h.integer();
                        if (nd < 0 || nd > 64) fail("string array has " + std::to_string(nd) + " dimensions");
                        for (std::int32_t a = 0; a < nd; ++a) {
                            const std::int32_t x = h.integer();
                            if (x < 0) fail("string array has dimension " + std::to_string(x));
                            shape.push_back(x);
                        }
                        filled = h.integer() != 0;
                    }
                    const std::size_t ne = product(shape);
                    const std::size_t head = d.fixed ? 0 : 4 * (shape.size() + 2);
                    if (filled && head + 4 * ne > std::size_t(std::max(ln, 0)))
                        fail("row " + fmt_u(row) + " of " + d.name + ": " + fmt_u(ne) + " strings do not fit in " +
                             std::to_string(ln) + " bytes");
                    if (!have) {
                        out.shape = shape;
                        have = true;
                    } else if (shape != out.shape) {
                        fail("rows of " + d.name + " differ in shape; read fewer rows at once");
                    }
                    if (filled)
                        h.strings(ne, out.strings);
                    else
                        out.strings.resize(out.strings.size() + ne);
                }
            });
            break;
        }
        case kIndirect: {
            std::vector<std::uint64_t> offs;
            visit(f, col, s, n, [&](const std::vector<std::uint8_t>& b, std::size_t o, std::size_t r, std::size_t k) {
                for (std::size_t i = 0; i < k; ++i) offs.push_back(u64_at(b, o + (r + i) * 8, big));
            });
            std::unique_ptr<ArrayFile> af;
            bool have = false;
            read_indirect(file_, big, d, offs, s, out, af, have);
            break;
        }
        }
        return out;
    }
};

// ------------------------------------------------------------ IncrementalStMan

// IncrementalStMan (core/src/ms.rs:1060-1348): values stored only where
// they change; per bucket and column, (first row, value offset) intervals
// in the bucket's index part; the bucket index after the initial buckets.
class Ism : public Manager {
public:
    Ism(const TableFiles& t, std::uint32_t seq, const std::vector<std::uint8_t>& blob,
        const std::vector<const ColumnDesc*>& cols)
        : file_(t.file(seq)) {
        Aips c(blob, true);
        c.magic();
        auto o = c.begin("ISM");
        (void)c.str();
        c.end(o);
        File f(file_);
        const auto head = f.at(0, kBucketStart);
        Aips h(head, !t.little);
        h.magic();
        auto ho = h.begin("IncrementalStMan");
        big_ = ho.version >= 5 ? h.boolean() : true;
        if (big_ == t.little) fail("IncrementalStMan endian flag mismatches the table");
        bucket_ = h.u32();
        const std::uint32_t ninit = h.u32();
        const std::uint64_t at = kBucketStart + std::uint64_t(ninit) * bucket_;
        const std::uint64_t size = f.size();
        if (at > size) fail("IncrementalStMan index beyond the file");
        const auto ix = f.at(at, std::size_t(size - at));
        Aips ic(ix, big_);
        ic.magic();
        auto io = ic.begin("ISMIndex");
        const std::uint32_t used = ic.u32();
        rows_ = ic.block(io.version > 1);
        auto bk = ic.block(false);
        rows_.resize(std::min<std::size_t>(used + 1, rows_.size()));
        bk.resize(std::min<std::size_t>(used, bk.size()));
        buckets_.assign(bk.begin(), bk.end());
        for (const ColumnDesc* d : cols) {
            Form m;
            m.dtype = d->dtype;
            if (!d->array) {
                m.kind = d->dtype == kString ? kStrs : kVals;
                m.n = 1;
            } else if (!d->direct()) {
                m.kind = kInd;
            } else if (!d->fixed || d->shape.empty()) {
                m.kind = kRefused;
                m.why = "IncrementalStMan Direct array column " + d->name + " without a fixed shape is not supported";
            } else {
                m.kind = d->dtype == kString ? kStrs : kVals;
                m.n = product(d->shape);
            }
            forms_.push_back(m);
        }
    }

    bool supports(int col, std::string* why) const override {
        if (forms_[col].kind != kRefused) return true;
        if (why) *why = forms_[col].why;
        return false;
    }

    Cells read(const ColumnDesc& d, int col, const std::vector<std::uint64_t>& rows, const Slice* cut) override {
        Cells acc;
        acc.dtype = d.dtype;
        bool first = true;
        for_runs(rows, [&](std::uint64_t s, std::uint64_t n) { merge(acc, first, run(d, col, s, n), d.name); });
        if (first) acc.nrows = 0;
        if (cut && acc.nrows > 0) return cut_cells(std::move(acc), *cut);
        return acc;
    }

private:
    enum Kind { kVals, kStrs, kInd, kRefused };
    struct Form {
        DType dtype;
        Kind kind;
        std::size_t n = 1;
        std::string why;
    };
    std::string file_;
    bool big_ = true;
    std::uint32_t bucket_ = 0;
    std::vector<std::uint64_t> rows_;
    std::vector<std::uint32_t> buckets_;
    std::vector<Form> forms_;

    // (first row in the bucket, value offset) of column `col`.
    std::vector<std::pair<std::uint64_t, std::uint32_t>> intervals(const std::vector<std::uint8_t>& b, int col) const {
        const std::uint32_t w = u32_at(b, 0, big_);
        const bool wide = (w & 0xf0000000u) != 0;
        std::size_t at = w & 0x0fffffffu;
        for (std::size_t c = 0; c < forms_.size(); ++c) {
            const std::uint32_t n = u32_at(b, at, big_);
            at += 4;
            std::vector<std::pair<std::uint64_t, std::uint32_t>> out(n);
            for (auto& e : out) {
                e.first = wide ? u64_at(b, at, big_) : u32_at(b, at, big_);
                at += wide ? 8 : 4;
            }
            for (auto& e : out) {
                e.second = u32_at(b, at, big_);
                at += 4;
            }
            if (int(c) == col) {
                if (out.empty()) fail("IncrementalStMan bucket without values for the column");
                return out;
            }
        }
        fail("IncrementalStMan bucket without values for the column");
    }

    // The strings of one value: its total length (itself included), then
    // one string's characters, or per string a length and characters.
    std::vector<std::string> strings_of(const std::uint8_t* p, std::size_t avail, std::size_t n) const {
        if (avail < 4) fail("truncated IncrementalStMan string");
        const std::size_t total = get32(p, big_);
        if (total > avail || total < 4) fail("truncated IncrementalStMan string");
        std::vector<std::string> out;
        if (n == 1) {
            out.push_back(text(p + 4, total - 4));
            return out;
        }
        std::size_t at = 4;
        for (std::size_t i = 0; i < n; ++i) {
            if (at + 4 > total) fail("truncated IncrementalStMan string");
            const std::size_t len = get32(p + at, big_);
            if (at + 4 + len > total) fail("truncated IncrementalStMan string");
            out.push_back(text(p + at + 4, len));
            at += 4 + len;
        }
        return out;
    }

    Cells run(const ColumnDesc& d, int col, std::uint64_t s, std::uint64_t n) const {
        const Form& m = forms_[col];
        if (m.kind == kRefused) fail(m.why);
        File f(file_);
        Cells out;
        out.dtype = d.dtype;
        out.nrows = n;
        if (d.array && m.kind != kInd) out.shape = d.shape;
        std::unique_ptr<ArrayFile> af;
        bool have = false;
        std::uint64_t row = s;
        while (row < s + n) {
            std::size_t i = std::size_t(std::upper_bound(rows_.begin(), rows_.end(), row) - rows_.begin());
            i = i == 0 ? 0 : i - 1;
            if (i + 1 >= rows_.size()) fail("row " + fmt_u(row) + " beyond the IncrementalStMan index");
            const std::uint64_t b0 = rows_[i], b1 = rows_[i + 1];
            const auto b = f.at(kBucketStart + std::uint64_t(buckets_[i]) * bucket_, bucket_);
            const auto iv = intervals(b, col);
            const std::uint64_t last = std::min(b1 - 1, s + n - 1);
            while (row <= last) {
                const std::uint64_t local = row - b0;
                std::size_t k = std::size_t(std::upper_bound(iv.begin(), iv.end(), std::make_pair(local, ~0u)) - iv.begin());
                k = k == 0 ? 0 : k - 1;
                const std::uint64_t until = k + 1 < iv.size() ? iv[k + 1].first : b1 - b0;
                const std::size_t cnt = std::size_t(std::min(until, last - b0 + 1) - local);
                const std::size_t at = 4 + std::size_t(iv[k].second);
                const std::uint8_t* p = at <= b.size() ? b.data() + at : b.data() + b.size();
                const std::size_t avail = at <= b.size() ? b.size() - at : 0;
                if (m.kind == kVals) {
                    std::vector<std::uint8_t> one;
                    if (m.dtype == kBool) {
                        if (m.n == 1)
                            one.push_back(avail ? (p[0] & 1) : 0);
                        else
                            append_bits(one, p, avail, 0, m.n);
                    } else {
                        append_values(one, p, avail, m.dtype, m.n, big_);
                    }
                    for (std::size_t r = 0; r < cnt; ++r) out.bytes.insert(out.bytes.end(), one.begin(), one.end());
                } else if (m.kind == kStrs) {
                    const auto one = strings_of(p, avail, m.n);
                    for (std::size_t r = 0; r < cnt; ++r) out.strings.insert(out.strings.end(), one.begin(), one.end());
                } else {
                    const std::uint64_t off = u64_at(b, at, big_);
                    if (off == 0) fail("row " + fmt_u(row) + " of " + d.name + " has no array");
                    Cells one;
                    std::vector<std::uint64_t> offs{off};
                    bool h1 = false;
                    read_indirect(file_, big_, d, offs, row, one, af, h1);
                    if (!have) {
                        out.shape = one.shape;
                        have = true;
                    } else if (one.shape != out.shape) {
                        fail("rows of " + d.name + " differ in shape; read fewer rows at once");
                    }
                    for (std::size_t r = 0; r < cnt; ++r) {
                        out.bytes.insert(out.bytes.end(), one.bytes.begin(), one.bytes.end());
                        out.strings.insert(out.strings.end(), one.strings.begin(), one.strings.end());
                    }
                }
                row += cnt;
            }
        }
        return out;
    }
};

// ------------------------------------------------------------ tiled managers

// TiledShapeStMan, TiledColumnStMan and TiledCellStMan
// (core/src/ms.rs:1329-1727): hypercubes of tiles, one bucket per tile in a
// _TSM file, the data columns sharing a tile in descending order of value
// size. A TiledShapeStMan maps rows to (cube, position on the last axis);
// a TiledColumnStMan has one cube whose last axis is the row; a
// TiledCellStMan has one cube per row (read with a trailing axis of 1).
class Tsm : public Manager {
public:
    Tsm(const TableFiles& t, std::uint32_t seq, const std::string& kind, const std::vector<const ColumnDesc*>& cols) {
        const auto bytes = slurp(t.file(seq));
        Aips h(bytes, true);
        h.magic();
        auto outer = h.begin(kind.c_str());
        const bool fixed_tile = kind == "TiledColumnStMan" || kind == "TiledCellStMan";
        if (fixed_tile) (void)h.ipos();
        auto inner = h.begin("TiledStMan");
        big_ = inner.version >= 2 ? h.boolean() : true;
        if (big_ == t.little) fail("TiledStMan endian flag mismatches the table");
        const std::uint32_t sq = h.u32();
        const std::uint64_t nrow = inner.version >= 3 ? h.u64() : h.u32();
        const std::uint32_t ncol = h.u32();
        if (sq != seq || ncol != cols.size() || nrow != t.nrows)
            fail("TiledStMan header of " + t.file(seq) + " mismatches the table");
        for (const ColumnDesc* d : cols) {
            if (h.i32() != int(d->dtype)) fail("TiledStMan header data type mismatches the column");
            dtypes_.push_back(d->dtype);
        }
        (void)h.str();  // hypercolumn
        if (inner.version >= 3)
            (void)h.u64();
        else
            (void)h.u32();  // cache size
        per_row_ = kind == "TiledCellStMan";
        nrdim_ = int(h.u32()) + (per_row_ ? 1 : 0);
        const std::uint64_t nfile = inner.version >= 3 ? h.u64() : h.u32();
        for (std::uint64_t i = 0; i < nfile; ++i) {
            if (h.boolean()) {
                const std::uint32_t v = h.u32();
                const std::uint32_t fs = h.u32();
                if (v == 1)
                    (void)h.u32();
                else
                    (void)h.u64();
                files_.push_back(t.file(seq, "_TSM" + std::to_string(fs)));
            } else {
                files_.push_back("");
            }
        }
        const std::uint64_t ncube = inner.version >= 3 ? h.u64() : h.u32();
        for (std::uint64_t i = 0; i < ncube; ++i) {
            Cube c;
            const std::uint32_t v = h.u32();
            h.skip();              // coordinates record
            (void)h.boolean();     // extensible
            (void)h.u32();         // dimensions
            c.shape = h.ipos();
            c.tile = h.ipos();
            if (per_row_ && !c.shape.empty()) {
                c.shape.push_back(1);
                c.tile.push_back(1);
            }
            c.file = h.i32();
            c.offset = v == 1 ? h.u32() : h.u64();
            c.tiles.resize(c.shape.size());
            for (std::size_t a = 0; a < c.shape.size(); ++a)
                c.tiles[a] = c.tile[a] > 0 ? (c.shape[a] + c.tile[a] - 1) / c.tile[a] : 0;
            c.pixels = product(c.tile);
            c.bucket = 0;
            for (DType d : dtypes_) c.bucket += tile_bytes(d, c.pixels);
            cubes_.push_back(std::move(c));
        }
        h.end(inner);
        by_row_ = !fixed_tile;
        if (by_row_) {
            (void)h.ipos();  // default tile
            const std::uint32_t n = h.u32();
            auto rm = h.block(false), cm = h.block(false), pm = h.block(false);
            rm.resize(std::min<std::size_t>(n, rm.size()));
            cm.resize(std::min<std::size_t>(n, cm.size()));
            pm.resize(std::min<std::size_t>(n, pm.size()));
            row_map_.assign(rm.begin(), rm.end());
            cube_map_.assign(cm.begin(), cm.end());
            pos_map_.assign(pm.begin(), pm.end());
        }
        h.end(outer);
        // stable order by descending value size
        order_.resize(dtypes_.size());
        for (std::size_t i = 0; i < order_.size(); ++i) order_[i] = int(i);
        std::stable_sort(order_.begin(), order_.end(),
                         [&](int a, int b) { return dtype_size(dtypes_[a]) > dtype_size(dtypes_[b]); });
    }

    Cells read(const ColumnDesc& d, int col, const std::vector<std::uint64_t>& rows, const Slice* cut) override {
        const DType dt = dtypes_[col];
        if (nrdim_ < 2) fail("tiled scalar column " + d.name + " is not supported");
        Cells out;
        out.dtype = dt;
        out.nrows = rows.size();
        const int last = nrdim_ - 1;
        const std::size_t ns = native_size(dt);
        std::vector<std::int64_t> cell0, lo;
        bool have = false;
        std::map<std::pair<std::size_t, std::uint64_t>, std::vector<std::uint8_t>> cache;
        std::map<int, std::unique_ptr<File>> handles;
        for (std::uint64_t row : rows) {
            std::size_t ci;
            std::int64_t pos;
            locate(row, ci, pos);
            const Cube& c = cubes_[ci];
            const std::vector<std::int64_t> cell(c.shape.begin(), c.shape.begin() + last);
            if (!have) {
                Slice reg;
                if (cut) {
                    reg = resolve_slice(*cut, cell);
                } else {
                    for (auto n : cell) reg.push_back({0, n});
                }
                lo.clear();
                out.shape.clear();
                for (auto& r : reg) {
                    lo.push_back(r.first);
                    out.shape.push_back(r.second);
                }
                cell0 = cell;
                have = true;
            } else if (cell != cell0) {
                fail("rows of " + d.name + " differ in shape; read fewer rows at once");
            }
            if (c.file < 0 || std::size_t(c.file) >= files_.size() || files_[c.file].empty())
                fail("hypercube of " + d.name + " has no file");
            auto& fh = handles[c.file];
            if (!fh) fh.reset(new File(files_[c.file]));
            const std::size_t col_at = offset_in_tile(col, c);
            const std::int64_t tlast = pos / c.tile[last], in_last = pos % c.tile[last];
            std::vector<std::int64_t> gstride(nrdim_, 1), pstride(nrdim_, 1);
            for (int a = 1; a < nrdim_; ++a) {
                gstride[a] = gstride[a - 1] * c.tiles[a - 1];
                pstride[a] = pstride[a - 1] * c.tile[a - 1];
            }
            const std::vector<std::int64_t>& len = out.shape;
            const std::size_t npix = product(len);
            const std::size_t base_out = out.bytes.size();
            out.bytes.resize(base_out + npix * ns);
            // runs along axis 0 that stay within one tile
            std::vector<std::int64_t> idx(last, 0);  // position in the region, axes >= 1
            for (std::size_t e = 0; e < npix;) {
                // output pixel e: axis 0 coordinate
                const std::int64_t x0 = std::int64_t(e % std::size_t(len[0]));
                std::size_t rem = e / std::size_t(len[0]);
                std::uint64_t tile_nr = std::uint64_t(tlast) * std::uint64_t(gstride[last]);
                std::int64_t pix = in_last * pstride[last];
                for (int a = 1; a < last; ++a) {
                    const std::int64_t xa = lo[a] + std::int64_t(rem % std::size_t(len[a]));
                    rem /= std::size_t(len[a]);
                    tile_nr += std::uint64_t(xa / c.tile[a] * gstride[a]);
                    pix += (xa % c.tile[a]) * pstride[a];
                }
                const std::int64_t g0 = lo[0] + x0;
                tile_nr += std::uint64_t(g0 / c.tile[0]);
                pix += g0 % c.tile[0];
                const std::int64_t run = std::min<std::int64_t>(len[0] - x0, c.tile[0] - g0 % c.tile[0]);
                const auto key = std::make_pair(ci, tile_nr);
                auto it = cache.find(key);
                if (it == cache.end()) {
                    if (cache.size() >= kTileCache) cache.clear();
                    it = cache.emplace(key, fh->at(c.offset + tile_nr * c.bucket, c.bucket)).first;
                }
                const std::vector<std::uint8_t>& tb = it->second;
                std::uint8_t* dst = out.bytes.data() + base_out + e * ns;
                if (dt == kBool) {
                    std::vector<std::uint8_t> bits;
                    if (col_at > tb.size()) fail("tile overrun");
                    append_bits(bits, tb.data() + col_at, tb.size() - col_at, std::size_t(pix), std::size_t(run));
                    std::memcpy(dst, bits.data(), bits.size());
                } else {
                    const std::size_t sz = dtype_size(dt);
                    const std::size_t from = col_at + std::size_t(pix) * sz;
                    if (from + std::size_t(run) * sz > tb.size()) fail("tile overrun");
                    std::memcpy(dst, tb.data() + from, std::size_t(run) * sz);
                    if (big_ && sz > 1) {
                        const std::size_t unit = dt == kComplex ? 4 : dt == kDComplex ? 8 : sz;
                        for (std::size_t k = 0; k < std::size_t(run) * sz; k += unit) std::reverse(dst + k, dst + k + unit);
                    }
                }
                e += std::size_t(run);
            }
        }
        if (!have) {
            out.shape.clear();
        }
        return out;
    }

private:
    struct Cube {
        std::vector<std::int64_t> shape, tile, tiles;
        int file = -1;
        std::uint64_t offset = 0;
        std::size_t bucket = 0, pixels = 0;
    };
    static std::size_t tile_bytes(DType d, std::size_t pixels) {
        return d == kBool ? (pixels + 7) / 8 : dtype_size(d) * pixels;
    }
    bool big_ = true, per_row_ = false, by_row_ = false;
    int nrdim_ = 0;
    std::vector<DType> dtypes_;
    std::vector<std::string> files_;
    std::vector<Cube> cubes_;
    std::vector<std::uint32_t> row_map_, cube_map_, pos_map_;
    std::vector<int> order_;

    std::size_t offset_in_tile(int col, const Cube& c) const {
        std::size_t at = 0;
        for (int k : order_) {
            if (k == col) return at;
            at += tile_bytes(dtypes_[k], c.pixels);
        }
        return at;
    }

    void locate(std::uint64_t row, std::size_t& cube, std::int64_t& pos) const {
        if (per_row_) {
            if (row >= cubes_.size() || cubes_[row].shape.empty()) fail("row " + fmt_u(row) + " has no tiled array");
            cube = std::size_t(row);
            pos = 0;
            return;
        }
        if (!by_row_) {
            if (cubes_.empty()) fail("row " + fmt_u(row) + " has no tiled array");
            cube = 0;
            pos = std::int64_t(row);
            return;
        }
        if (row_map_.empty() || row > row_map_.back()) fail("row " + fmt_u(row) + " has no tiled array");
        const std::size_t i = std::size_t(std::lower_bound(row_map_.begin(), row_map_.end(), std::uint32_t(row)) - row_map_.begin());
        cube = cube_map_[i];
        if (cube == 0 || cube >= cubes_.size()) fail("row " + fmt_u(row) + " has no tiled array");
        pos = std::int64_t(pos_map_[i]) - (std::int64_t(row_map_[i]) - std::int64_t(row));
    }
};

// The row count of table.lock's sync record, when it has one
// (core/src/ms.rs:2248-2276): after 65 request words, a canonical length and
// an AipsIO "sync" object starting with the row count.
bool lock_rows(const std::string& dir, std::uint64_t& rows) {
    std::ifstream f(dir + "/table.lock", std::ios::binary);
    if (!f) return false;
    std::vector<std::uint8_t> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    constexpr std::size_t kReq = (1 + 2 * 32) * 4;
    if (b.size() < kReq + 4) return false;
    const std::size_t len = get32(b.data() + kReq, true);
    if (len == 0) return false;
    if (kReq + 4 + len > b.size()) fail("truncated table.lock sync record");
    std::vector<std::uint8_t> info(b.begin() + kReq + 4, b.begin() + kReq + 4 + len);
    Aips c(info, true);
    c.magic();
    auto o = c.begin("sync");
    rows = o.version == 1 ? c.u32() : c.u64();
    return true;
}

std::string json_str(const std::string& s) {
    std::string o = "\"";
    for (unsigned char ch : s) {
        if (ch == '"')
            o += "\\\"";
        else if (ch == '\\')
            o += "\\\\";
        else if (ch < 0x20) {
            char buf[8];
            std::snprintf(buf, sizeof buf, "\\u%04x", ch);
            o += buf;
        } else {
            o += char(ch);
        }
    }
    return o + "\"";
}

}  // namespace

DType dtype_of(int code) {
    switch (code) {
    case 0: case 1: case 2: case 3: case 4: case 5: case 6: case 7: case 8: case 9: case 10: case 11: case 29:
        return DType(code);
    default:
        fail("unsupported column data type code " + std::to_string(code));
    }
}

std::size_t dtype_size(DType d) {
    switch (d) {
    case kChar: case kUChar: return 1;
    case kShort: case kUShort: return 2;
    case kInt: case kUInt: case kFloat: return 4;
    case kDouble: case kInt64: case kComplex: return 8;
    case kDComplex: return 16;
    default: return 0;
    }
}

std::size_t native_size(DType d) { return d == kBool ? 1 : dtype_size(d); }

const char* numpy_name(DType d) {
    switch (d) {
    case kBool: return "bool";
    case kChar: return "int8";
    case kUChar: return "uint8";
    case kShort: return "int16";
    case kUShort: return "uint16";
    case kInt: return "int32";
    case kUInt: return "uint32";
    case kFloat: return "float32";
    case kDouble: return "float64";
    case kComplex: return "complex64";
    case kDComplex: return "complex128";
    case kString: return "str";
    case kInt64: return "int64";
    }
    return "?";
}

Slice resolve_slice(const Slice& cut, const std::vector<std::int64_t>& shape) {
    if (cut.empty() || cut.size() > shape.size())
        fail("a slice of " + fmt_u(cut.size()) + " axes for cells of " + fmt_u(shape.size()) + " axes");
    Slice full(shape.size() - cut.size(), {0, -1});
    full.insert(full.end(), cut.begin(), cut.end());
    Slice out;
    for (std::size_t i = 0; i < shape.size(); ++i) {
        const std::int64_t a = full[i].first, n = full[i].second < 0 ? shape[i] - a : full[i].second;
        if (a < 0 || n < 1 || a + n > shape[i])
            fail("slice " + std::to_string(a) + "+" + std::to_string(n) + " of axis " + fmt_u(i) +
                 " is outside a cell of " + std::to_string(shape[i]));
        out.push_back({a, n});
    }
    return out;
}

Table::~Table() = default;

// table.dat (core/src/ms.rs:1906-2060): the Table object (rows, endianness,
// "PlainTable"), the TableDesc with its column descriptions, the ColumnSet
// with the data managers and per-column bindings, then one header blob per
// manager.
Table::Table(const std::string& dir) : path(dir) {
    const auto bytes = slurp(dir + "/table.dat");
    Aips c(bytes, true);
    c.magic();
    auto tobj = c.begin("Table");
    if (tobj.version > 3) fail("table.dat version " + fmt_u(tobj.version) + " is not supported");
    nrows = tobj.version > 2 ? c.u64() : c.u32();
    std::uint64_t synced = 0;
    if (lock_rows(dir, synced) && synced > 0) nrows = synced;
    little = c.u32() != 0;
    const std::string kind = c.str();
    if (kind != "PlainTable") fail("table type " + kind + " is not supported");
    auto desc = c.begin("TableDesc");
    (void)c.str();
    (void)c.str();
    (void)c.str();
    c.skip();
    if (desc.version != 1) c.skip();
    const std::uint32_t ncol = c.u32();
    for (std::uint32_t i = 0; i < ncol; ++i) {
        ColumnDesc d;
        (void)c.u32();
        const std::string cls = c.str();
        (void)c.u32();
        d.name = c.str();
        (void)c.str();  // comment
        d.manager_type = c.str();
        d.manager_name = c.str();
        d.dtype = dtype_of(c.i32());
        d.option = c.i32();
        d.ndim = c.i32();
        d.array = cls.rfind("ArrayColumnDesc", 0) == 0;
        if (!d.array && cls.rfind("ScalarColumnDesc", 0) != 0)
            fail("column class " + cls + " of " + d.name + " is not supported");
        if (d.array) d.shape = c.ipos();
        d.max_length = c.u32();
        c.skip();  // keywords
        (void)c.u32();
        if (d.array) {
            (void)c.boolean();
        } else if (d.dtype == kString) {
            (void)c.str();  // default value
        } else if (d.dtype == kBool) {
            (void)c.u8();
        } else {
            (void)c.take(dtype_size(d.dtype));
        }
        d.fixed = !d.shape.empty() && (d.option & 4);
        columns.push_back(std::move(d));
    }
    c.end(desc);
    if (tobj.version == 1) c.skip();
    const std::int32_t first = c.i32();
    std::int32_t setv = 1;
    if (first < 0) {
        setv = -first;
        if (setv <= 2)
            (void)c.u32();
        else
            (void)c.u64();
    }
    if (setv >= 3) {
        (void)c.i32();  // storage option
        (void)c.i32();  // buffer size
    }
    (void)c.u32();
    const std::uint32_t nm = c.u32();
    std::vector<std::pair<std::string, std::uint32_t>> mans;
    for (std::uint32_t i = 0; i < nm; ++i) {
        std::string k = c.str();
        const std::uint32_t s = c.u32();
        mans.push_back({k, s});
    }
    for (auto& d : columns) {
        (void)c.u32();
        (void)c.str();
        (void)c.u32();
        d.manager_seq = c.u32();
        if (d.array && c.boolean()) {
            auto s = c.ipos();
            if (d.shape.empty()) d.shape = s;
            d.fixed = !d.shape.empty();
        }
    }
    std::vector<std::vector<std::uint8_t>> blobs;
    for (std::uint32_t i = 0; i < nm; ++i) {
        const std::uint32_t len = c.u32();
        const std::uint8_t* p = c.take(len);
        blobs.emplace_back(p, p + len);
    }
    const TableFiles tf{dir, little, nrows};
    binding_.assign(columns.size(), {-1, -1});
    for (std::uint32_t i = 0; i < nm; ++i) {
        Slot s;
        s.kind = mans[i].first;
        s.seq = mans[i].second;
        std::vector<const ColumnDesc*> mine;
        for (std::size_t k = 0; k < columns.size(); ++k)
            if (columns[k].manager_seq == s.seq) {
                binding_[k] = {int(i), int(mine.size())};
                mine.push_back(&columns[k]);
            }
        if (s.kind == "StandardStMan") {
            s.m.reset(new Ssm(tf, s.seq, blobs[i], mine));
        } else if (s.kind == "IncrementalStMan") {
            try {
                s.m.reset(new Ism(tf, s.seq, blobs[i], mine));
            } catch (const Error& e) {
                s.refused = e.what();
            }
        } else if (s.kind == "TiledShapeStMan" || s.kind == "TiledColumnStMan" || s.kind == "TiledCellStMan") {
            s.m.reset(new Tsm(tf, s.seq, s.kind, mine));
        } else {
            s.refused = "storage manager " + s.kind + " is not supported";
        }
        slots_.push_back(std::move(s));
    }
}

int Table::column(const std::string& name) const {
    for (std::size_t i = 0; i < columns.size(); ++i)
        if (columns[i].name == name) return int(i);
    fail("column " + name + " does not exist");
}

Cells Table::read_impl(int index, const std::vector<std::uint64_t>& rows, const Slice* cut) const {
    const ColumnDesc& d = columns[index];
    const auto b = binding_[index];
    if (b.first < 0) fail(d.name + ": column without a data manager");
    const Slot& s = slots_[b.first];
    try {
        if (!s.m) fail(s.refused);
        std::string why;
        if (!s.m->supports(b.second, &why)) fail(why);
        Cells out = s.m->read(d, b.second, rows, cut);
        out.dtype = d.dtype;
        if (rows.empty()) {
            out.nrows = 0;
            out.bytes.clear();
            out.strings.clear();
            if (cut && !d.shape.empty() && cut->size() <= d.shape.size()) {
                out.shape.clear();
                for (auto& r : resolve_slice(*cut, d.shape)) out.shape.push_back(r.second);
            } else {
                out.shape = d.shape;
            }
        }
        return out;
    } catch (const Error& e) {
        throw Error(d.name + ": " + e.what());
    }
}

Cells Table::read(const std::string& name, std::uint64_t start, std::uint64_t count, const Slice* cut) const {
    const int i = column(name);
    if (count == 0) count = nrows > start ? nrows - start : 0;
    if (start + count > nrows)
        fail("rows " + fmt_u(start) + ".." + fmt_u(start + count) + " exceed the " + fmt_u(nrows) + " rows of the table");
    std::vector<std::uint64_t> rows(count);
    for (std::uint64_t k = 0; k < count; ++k) rows[k] = start + k;
    return read_impl(i, rows, cut);
}

Cells Table::read_rows(const std::string& name, const std::vector<std::uint64_t>& rows, const Slice* cut) const {
    const int i = column(name);
    for (auto r : rows)
        if (r >= nrows) fail("row " + fmt_u(r) + " exceeds the " + fmt_u(nrows) + " rows of the table");
    return read_impl(i, rows, cut);
}

std::string Table::describe_json() const {
    std::ostringstream o;
    o << "{\"path\":" << json_str(path) << ",\"nrows\":" << nrows << ",\"little_endian\":" << (little ? "true" : "false")
      << ",\"columns\":[";
    for (std::size_t i = 0; i < columns.size(); ++i) {
        const ColumnDesc& c = columns[i];
        const auto b = binding_[i];
        bool ok = b.first >= 0 && slots_[b.first].m != nullptr;
        if (ok) ok = slots_[b.first].m->supports(b.second, nullptr);
        if (i) o << ',';
        o << "{\"name\":" << json_str(c.name) << ",\"dtype\":\"" << numpy_name(c.dtype) << "\",\"is_array\":"
          << (c.array ? "true" : "false") << ",\"option\":" << c.option << ",\"ndim\":" << c.ndim << ",\"shape\":[";
        for (std::size_t k = 0; k < c.shape.size(); ++k) o << (k ? "," : "") << c.shape[k];
        o << "],\"max_length\":" << c.max_length << ",\"manager_type\":"
          << json_str(b.first >= 0 ? slots_[b.first].kind : c.manager_type) << ",\"manager_name\":"
          << json_str(c.manager_name) << ",\"manager_seq\":" << c.manager_seq << ",\"supported\":"
          << (ok ? "true" : "false") << "}";
    }
    o << "]}";
    return o.str();
}

}  // namespace ms
}  // namespace rang
