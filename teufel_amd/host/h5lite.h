// h5lite.h - a minimal read-only HDF5 reader for TEUFEL's screen files.
//
// TEUFEL stores meshed screens in HDF5 (src/mesh.cpp:30-218 reads them with
// libhdf5: datasets MeshCornerPoints, MeshTriangles, ObservationPosition,
// ObservationTime, ElMagField, each with integer/double attributes).  This
// reader covers the subset of the HDF5 file format (version 0/1 superblock,
// 8-byte offsets and lengths, symbol-table root group, version 1 object
// headers, contiguous or compact layouts, little-endian integer and IEEE
// float data, version 1-3 attributes) that such files use, without linking
// libhdf5.  Anything outside that subset raises IOexception naming it.
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
