// h5lite.h - a minimal HDF5 reader and writer for TEUFEL's screen files.
//
// TEUFEL stores meshed screens in HDF5 (src/mesh.cpp:30-218 reads them with
// libhdf5: datasets MeshCornerPoints, MeshTriangles, ObservationPosition,
// ObservationTime, ElMagField, each with integer/double attributes).  This
// reader covers the subset of the HDF5 file format (version 0/1 superblock,
// 8-byte offsets and lengths, symbol-table root group, version 1 object
// headers, contiguous or compact layouts, little-endian integer and IEEE
// float data, version 1-3 attributes) that such files use, without linking
// libhdf5.  Anything outside that subset raises IOexception naming it.
// H5LiteWriter writes files of the same subset (superblock version 0, one
// root group of up to eight contiguous datasets with scalar attributes), the
// layout MeshedScreen::writeFile produces with libhdf5 (src/mesh.cpp:427-736).
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

class H5LiteFile {
public:
    explicit H5LiteFile(const std::string& filename);

    bool exists(const std::string& dataset) const;
    // dimensions of a dataset (row-major, slowest first)
    std::vector<uint64_t> dims(const std::string& dataset) const;
    // all elements of a dataset converted to double / to int64
    std::vector<double> read_double(const std::string& dataset) const;
    std::vector<int64_t> read_int(const std::string& dataset) const;
    // a scalar (or first element of a) numeric attribute of a dataset
    double attr_double(const std::string& dataset, const std::string& name) const;
    int64_t attr_int(const std::string& dataset, const std::string& name) const;

    struct Type {
        int cls = -1;          // 0 fixed point, 1 floating point
        uint32_t size = 0;     // bytes per element
        bool is_signed = true;
    };
    struct Attribute {
        Type type;
        uint64_t count = 0;
        std::vector<uint8_t> data;
    };
    struct Dataset {
        std::vector<uint64_t> dims;
        Type type;
        int layout = -1;          // 0 compact, 1 contiguous
        uint64_t address = 0;     // contiguous data
        uint64_t size = 0;
        std::vector<uint8_t> compact;
        std::map<std::string, Attribute> attrs;
        uint64_t sym_btree = ~0ull, sym_heap = ~0ull;  // symbol table message (groups)
    };

private:
    std::string fname;
    std::vector<uint8_t> buf;
    std::map<std::string, Dataset> sets;

    const Dataset& get(const std::string& name) const;
    void read_group(uint64_t btree, uint64_t heap);
    Dataset read_object(uint64_t header) const;
    [[noreturn]] void fail(const std::string& what) const;
    void need(uint64_t off, uint64_t n) const;
    uint64_t u(uint64_t off, int bytes) const;
};

class H5LiteWriter {
public:
    // dims row-major (slowest first); data has prod(dims) elements
    void add_double(const std::string& name, const std::vector<uint64_t>& dims, const std::vector<double>& data);
    void add_int32(const std::string& name, const std::vector<uint64_t>& dims, const std::vector<int32_t>& data);
    // scalar attributes of an added dataset
    void attr_int32(const std::string& dataset, const std::string& name, int32_t value);
    void attr_double(const std::string& dataset, const std::string& name, double value);
    void write(const std::string& filename) const;

private:
    struct Attr {
        std::string name;
        bool is_double;
        double d;
        int32_t i;
    };
    struct Set {
        std::string name;
        std::vector<uint64_t> dims;
        bool is_double;
        std::vector<uint8_t> bytes;
        std::vector<Attr> attrs;
    };
    std::vector<Set> sets;
    Set& find(const std::string& name);
};

