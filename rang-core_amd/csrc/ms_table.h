// Read-only access to casacore tables (Measurement Sets) for ingest to the
// device: table.dat, and the StandardStMan, IncrementalStMan, TiledShapeStMan,
// TiledColumnStMan and TiledCellStMan data files, with the scope and the
// refusals of rang-core's reader (core/src/ms.rs:1-18): scalar, direct-array
// and indirect-array columns of the numeric, boolean and string types;
// other storage managers, MultiFile tables and big-endian hosts refused.
// Cells come back as native bytes in casacore's (Fortran) axis order,
// booleans one byte each, strings as a list.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace rang {
namespace ms {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// casacore DataType codes a column can hold (core/src/ms.rs:57-105).
enum DType : int {
    kBool = 0,
    kChar = 1,
    kUChar = 2,
    kShort = 3,
    kUShort = 4,
    kInt = 5,
    kUInt = 6,
    kFloat = 7,
    kDouble = 8,
    kComplex = 9,
    kDComplex = 10,
    kString = 11,
    kInt64 = 29
};

DType dtype_of(int code);
// Bytes of one stored value (0 for booleans, which are bits, and strings).
std::size_t dtype_size(DType d);
// Bytes of one returned value (booleans one byte).
std::size_t native_size(DType d);
const char* numpy_name(DType d);

// A column of table.dat with its data-manager binding.
struct ColumnDesc {
    std::string name;
    DType dtype = kInt;
    bool array = false;
    int option = 0;  // 1 Direct, 2 Undefined, 4 FixedShape
    int ndim = 0;
    std::vector<std::int64_t> shape;  // casacore order; empty if not fixed
    bool fixed = false;
    std::uint32_t max_length = 0;
    std::string manager_type, manager_name;
    std::uint32_t manager_seq = 0;
    bool direct() const { return option & 1; }
};

// The cells of one read: `nrows` cells of `shape` as native bytes, or
// strings for string columns.
struct Cells {
    DType dtype = kInt;
    std::vector<std::int64_t> shape;
    std::uint64_t nrows = 0;
    std::vector<std::uint8_t> bytes;
    std::vector<std::string> strings;
};

// A cell region per axis (casacore order): first pixel and length, a
// negative length running to the end of the axis.
using Slice = std::vector<std::pair<std::int64_t, std::int64_t>>;

class Manager {
public:
    virtual ~Manager() = default;
    // Cells of rows `rows` (any order) of the manager's column `col`.
    virtual Cells read(const ColumnDesc& d, int col, const std::vector<std::uint64_t>& rows, const Slice* cut) = 0;
    // Whether column `col` reads; `why` the refusal otherwise.
    virtual bool supports(int col, std::string* why) const {
        (void)col;
        (void)why;
        return true;
    }
};

class Table {
public:
    explicit Table(const std::string& dir);
    ~Table();
    std::string path;
    std::uint64_t nrows = 0;
    bool little = true;
    std::vector<ColumnDesc> columns;

    int column(const std::string& name) const;  // throws when absent
    // Rows start .. start + count (count 0: to the end).
    Cells read(const std::string& name, std::uint64_t start, std::uint64_t count, const Slice* cut) const;
    // The rows `rows` in that order.
    Cells read_rows(const std::string& name, const std::vector<std::uint64_t>& rows, const Slice* cut) const;
    std::string describe_json() const;

private:
    struct Slot {
        std::string kind;
        std::uint32_t seq = 0;
        std::unique_ptr<Manager> m;
        std::string refused;  // set when the manager itself is refused
    };
    std::vector<Slot> slots_;
    std::vector<std::pair<int, int>> binding_;  // (slot, column of the manager)
    Cells read_impl(int index, const std::vector<std::uint64_t>& rows, const Slice* cut) const;
};

// Region of a cell of `shape` that `cut` selects, lengths resolved.
Slice resolve_slice(const Slice& cut, const std::vector<std::int64_t>& shape);

}  // namespace ms
}  // namespace rang
