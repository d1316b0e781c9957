// h5lite.cpp - minimal HDF5 reader (format subset described in h5lite.h),
// written from the HDF5 file format specification: superblock, local heap,
// version 1 B-tree (group nodes) and symbol table nodes, version 1 object
// headers with dataspace (0x0001), datatype (0x0003), data layout (0x0008),
// attribute (0x000C), continuation (0x0010) and symbol table (0x0011)
// messages.
#include "h5lite.h"

#include <cstdio>
#include <cstring>

#include "global.h"

namespace {
const uint8_t kSignature[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
constexpr uint64_t kUndefined = ~0ull;

uint64_t pad8(uint64_t n) { return (n + 7) & ~7ull; }
}  // namespace

void H5LiteFile::fail(const std::string& what) const { throw IOexception("h5lite: " + fname + ": " + what); }

void H5LiteFile::need(uint64_t off, uint64_t n) const
{
    if (off > buf.size() || n > buf.size() - off) fail("truncated or corrupt file (read past the end)");
}

uint64_t H5LiteFile::u(uint64_t off, int bytes) const
{
    need(off, (uint64_t)bytes);
    uint64_t v = 0;
    for (int i = bytes - 1; i >= 0; --i) v = (v << 8) | buf[off + i];
    return v;
}

H5LiteFile::H5LiteFile(const std::string& filename) : fname(filename)
{
    FILE* f = std::fopen(filename.c_str(), "rb");
    if (!f) throw IOexception("h5lite: cannot open " + filename);
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(f);
        fail("cannot size the file");
    }
    buf.resize((size_t)n);
    size_t got = n ? std::fread(buf.data(), 1, (size_t)n, f) : 0;
    std::fclose(f);
    if (got != (size_t)n) fail("read error");

    // the superblock sits at 0 or after a user block of 512 * 2^k bytes
    uint64_t sb = kUndefined;
    for (uint64_t off = 0; off + 8 <= buf.size(); off = off ? off * 2 : 512)
        if (std::memcmp(&buf[off], kSignature, 8) == 0) {
            sb = off;
            break;
        }
    if (sb == kUndefined) fail("not an HDF5 file");
    const int version = (int)u(sb + 8, 1);
    if (version > 1) fail("superblock version " + std::to_string(version) + " is not supported");
    if (u(sb + 13, 1) != 8 || u(sb + 14, 1) != 8) fail("only 8-byte offsets and lengths are supported");
    uint64_t p = sb + (version == 0 ? 24 : 28);
    const uint64_t base = u(p, 8);
    if (base != sb) fail("base address other than the superblock is not supported");
    p += 32;  // base, free-space, end-of-file, driver-information addresses
    // root group symbol table entry
    const uint64_t root_header = u(p + 8, 8);
    const uint32_t cache = (uint32_t)u(p + 16, 4);
    uint64_t btree = kUndefined, heap = kUndefined;
    if (cache == 1) {
        btree = u(p + 24, 8);
        heap = u(p + 32, 8);
    } else {
        // the symbol table message of the root object header
        const Dataset root = read_object(root_header);
        btree = root.sym_btree;
        heap = root.sym_heap;
    }
    if (btree == kUndefined || heap == kUndefined) fail("root group has no symbol table");
    read_group(btree, heap);
}

void H5LiteFile::read_group(uint64_t btree, uint64_t heap)
{
    need(heap, 32);
    if (std::memcmp(&buf[heap], "HEAP", 4) != 0) fail("local heap signature missing");
    const uint64_t heap_size = u(heap + 8, 8);
    const uint64_t heap_data = u(heap + 24, 8);
    need(heap_data, heap_size);
    auto name_at = [&](uint64_t off) {
        if (off >= heap_size) fail("link name outside the local heap");
        const char* s = reinterpret_cast<const char*>(&buf[heap_data + off]);
        size_t len = strnlen(s, (size_t)(heap_size - off));
        return std::string(s, len);
    };
    std::vector<uint64_t> nodes(1, btree);
    int guard = 0;
    while (!nodes.empty()) {
        if (++guard > 100000) fail("group B-tree too large or cyclic");
        const uint64_t node = nodes.back();
        nodes.pop_back();
        need(node, 24);
        if (std::memcmp(&buf[node], "TREE", 4) == 0) {
            if (u(node + 4, 1) != 0) fail("B-tree node is not a group node");
            const uint64_t used = u(node + 6, 2);
            uint64_t q = node + 24;  // signature, type, level, entries, two siblings
            for (uint64_t i = 0; i < used; ++i) {
                nodes.push_back(u(q + 8, 8));  // key (heap offset), then child: a node or a symbol node
                q += 16;
            }
        } else if (std::memcmp(&buf[node], "SNOD", 4) == 0) {
            const uint64_t count = u(node + 6, 2);
            for (uint64_t i = 0; i < count; ++i) {
                const uint64_t e = node + 8 + 40 * i;
                const std::string name = name_at(u(e, 8));
                const uint64_t header = u(e + 8, 8);
                const uint32_t cache = (uint32_t)u(e + 16, 4);
                if (cache == 1) continue;  // a subgroup: screens keep their datasets at the root
                sets[name] = read_object(header);
            }
        } else {
            fail("group node signature missing");
        }
    }
}

H5LiteFile::Dataset H5LiteFile::read_object(uint64_t header) const
{
    need(header, 16);
    if (std::memcmp(&buf[header], "OHDR", 4) == 0) fail("version 2 object headers are not supported");
    if (u(header, 1) != 1) fail("object header version is not 1");
    Dataset ds;
    bool have_space = false, have_type = false;
    std::vector<std::pair<uint64_t, uint64_t>> blocks(1, {header + 16, u(header + 8, 4)});
    auto read_type = [&](uint64_t q, Type& t) {
        const uint64_t b0 = u(q, 1);
        t.cls = (int)(b0 & 0x0f);
        t.size = (uint32_t)u(q + 4, 4);
        const uint64_t bits = u(q + 1, 1);
        if ((t.cls == 0 || t.cls == 1) && (bits & 1)) fail("big-endian data is not supported");
        t.is_signed = t.cls == 0 ? ((bits & 0x08) != 0) : true;
    };
    auto read_space = [&](uint64_t q, std::vector<uint64_t>& dims) {
        const uint64_t ver = u(q, 1), rank = u(q + 1, 1);
        uint64_t d = q + (ver == 1 ? 8 : 4);
        if (ver != 1 && ver != 2) fail("dataspace message version " + std::to_string(ver));
        dims.clear();
        if (ver == 2 && u(q + 3, 1) == 2) {
            dims.push_back(0);  // null dataspace: no elements
            return;
        }
        for (uint64_t i = 0; i < rank; ++i) dims.push_back(u(d + 8 * i, 8));
    };
    for (size_t bi = 0; bi < blocks.size(); ++bi) {
        uint64_t q = blocks[bi].first;
        const uint64_t end = q + blocks[bi].second;
        need(q, blocks[bi].second);
        while (q + 8 <= end) {
            const uint64_t type = u(q, 2), size = u(q + 2, 2), flags = u(q + 4, 1);
            const uint64_t m = q + 8;
            need(m, size);
            if ((flags & 0x02) && (type == 0x03 || type == 0x01)) fail("shared datatype/dataspace messages are not supported");
            switch (type) {
            case 0x01:
                read_space(m, ds.dims);
                have_space = true;
                break;
            case 0x03:
                read_type(m, ds.type);
                have_type = true;
                break;
            case 0x08: {
                const uint64_t ver = u(m, 1);
                if (ver == 3) {
                    ds.layout = (int)u(m + 1, 1);
                    if (ds.layout == 0) {
                        const uint64_t n = u(m + 2, 2);
                        need(m + 4, n);
                        ds.compact.assign(buf.begin() + (long)(m + 4), buf.begin() + (long)(m + 4 + n));
                    } else if (ds.layout == 1) {
                        ds.address = u(m + 2, 8);
                        ds.size = u(m + 10, 8);
                    } else {
                        fail("chunked or virtual dataset layouts are not supported");
                    }
                } else if (ver == 1 || ver == 2) {
                    const uint64_t rank = u(m + 1, 1);
                    ds.layout = (int)u(m + 2, 1);
                    uint64_t d = m + 8;
                    if (ds.layout == 1) {
                        ds.address = u(d, 8);
                        d += 8;
                        ds.size = kUndefined;  // from dataspace x type below
                    } else if (ds.layout == 0) {
                        d += 4 * rank;
                        const uint64_t n = u(d, 4);
                        need(d + 4, n);
                        ds.compact.assign(buf.begin() + (long)(d + 4), buf.begin() + (long)(d + 4 + n));
                    } else {
                        fail("chunked dataset layouts are not supported");
                    }
                } else {
                    fail("data layout message version " + std::to_string(ver));
                }
                break;
            }
            case 0x0C: {
                const uint64_t ver = u(m, 1);
                if (ver < 1 || ver > 3) fail("attribute message version " + std::to_string(ver));
                const uint64_t nlen = u(m + 2, 2), tlen = u(m + 4, 2), slen = u(m + 6, 2);
                uint64_t d = m + (ver == 3 ? 9 : 8);
                need(d, nlen);
                std::string name(reinterpret_cast<const char*>(&buf[d]), strnlen(reinterpret_cast<const char*>(&buf[d]), nlen));
                d += ver == 1 ? pad8(nlen) : nlen;
                Attribute a;
                read_type(d, a.type);
                d += ver == 1 ? pad8(tlen) : tlen;
                std::vector<uint64_t> adims;
                read_space(d, adims);
                d += ver == 1 ? pad8(slen) : slen;
                a.count = 1;
                for (uint64_t x : adims) a.count *= x;
                const uint64_t nbytes = a.count * a.type.size;
                need(d, nbytes);
                a.data.assign(buf.begin() + (long)d, buf.begin() + (long)(d + nbytes));
                ds.attrs[name] = a;
                break;
            }
            case 0x11:
                ds.sym_btree = u(m, 8);
                ds.sym_heap = u(m + 8, 8);
                break;
            case 0x10:
                if (blocks.size() > 1000) fail("too many header continuation blocks");
                blocks.push_back({u(m, 8), u(m + 8, 8)});
                break;
            default:
                break;
            }
            q = m + size;
        }
    }
    if (!have_space || !have_type) ds.layout = -1;  // not a dataset (e.g. a group object)
    if (ds.layout == 1 && ds.size == kUndefined) {
        uint64_t n = ds.type.size;
        for (uint64_t x : ds.dims) n *= x;
        ds.size = n;
    }
    return ds;
}

const H5LiteFile::Dataset& H5LiteFile::get(const std::string& name) const
{
    auto it = sets.find(name);
    if (it == sets.end() || it->second.layout < 0) fail("no dataset " + name);
    return it->second;
}

bool H5LiteFile::exists(const std::string& dataset) const
{
    auto it = sets.find(dataset);
    return it != sets.end() && it->second.layout >= 0;
}

std::vector<uint64_t> H5LiteFile::dims(const std::string& dataset) const { return get(dataset).dims; }

namespace {
double element(const H5LiteFile::Type& t, const uint8_t* p)
{
    if (t.cls == 1 && t.size == 8) {
        double v;
        std::memcpy(&v, p, 8);
        return v;
    }
    if (t.cls == 1 && t.size == 4) {
        float v;
        std::memcpy(&v, p, 4);
        return v;
    }
    uint64_t raw = 0;
    for (int i = (int)t.size - 1; i >= 0; --i) raw = (raw << 8) | p[i];
    if (t.is_signed && t.size < 8 && (raw >> (8 * t.size - 1)) & 1) raw |= ~0ull << (8 * t.size);
    return t.is_signed ? (double)(int64_t)raw : (double)raw;
}

int64_t element_int(const H5LiteFile::Type& t, const uint8_t* p)
{
    if (t.cls == 1) return (int64_t)element(t, p);
    uint64_t raw = 0;
    for (int i = (int)t.size - 1; i >= 0; --i) raw = (raw << 8) | p[i];
    if (t.is_signed && t.size < 8 && (raw >> (8 * t.size - 1)) & 1) raw |= ~0ull << (8 * t.size);
    return (int64_t)raw;
}

uint64_t product(const std::vector<uint64_t>& dims, uint64_t size, bool& overflow)
{
    uint64_t n = size;
    overflow = false;
    for (uint64_t x : dims) {
        if (x != 0 && n > ~0ull / x) overflow = true;
        n *= x;
    }
    return n;
}

bool numeric(const H5LiteFile::Type& t)
{
    return (t.cls == 1 && (t.size == 4 || t.size == 8)) || (t.cls == 0 && t.size >= 1 && t.size <= 8);
}
}  // namespace

std::vector<double> H5LiteFile::read_double(const std::string& dataset) const
{
    const Dataset& ds = get(dataset);
    if (!numeric(ds.type)) fail(dataset + ": not an integer or IEEE float dataset");
    bool overflow;
    const uint64_t n = product(ds.dims, 1, overflow);
    const uint64_t bytes = product(ds.dims, ds.type.size, overflow);
    if (overflow) fail(dataset + ": dataspace too large");
    const uint8_t* src;
    if (ds.layout == 0) {
        if (ds.compact.size() < bytes) fail(dataset + ": compact data too short");
        src = ds.compact.data();
    } else {
        if (ds.address == kUndefined) {
            if (n) fail(dataset + ": no data written");
            return {};
        }
        if (ds.size < bytes) fail(dataset + ": stored size smaller than its dataspace");
        need(ds.address, bytes);
        src = &buf[ds.address];
    }
    std::vector<double> out(n);
    for (uint64_t i = 0; i < n; ++i) out[i] = element(ds.type, src + i * ds.type.size);
    return out;
}

std::vector<int64_t> H5LiteFile::read_int(const std::string& dataset) const
{
    const Dataset& ds = get(dataset);
    if (!numeric(ds.type)) fail(dataset + ": not an integer or IEEE float dataset");
    bool overflow;
    const uint64_t n = product(ds.dims, 1, overflow);
    const uint64_t bytes = product(ds.dims, ds.type.size, overflow);
    if (overflow) fail(dataset + ": dataspace too large");
    const uint8_t* src;
    if (ds.layout == 0) {
        if (ds.compact.size() < bytes) fail(dataset + ": compact data too short");
        src = ds.compact.data();
    } else {
        if (ds.size < bytes) fail(dataset + ": stored size smaller than its dataspace");
        need(ds.address, bytes);
        src = &buf[ds.address];
    }
    std::vector<int64_t> out(n);
    for (uint64_t i = 0; i < n; ++i) out[i] = element_int(ds.type, src + i * ds.type.size);
    return out;
}

double H5LiteFile::attr_double(const std::string& dataset, const std::string& name) const
{
    const Dataset& ds = get(dataset);
    auto it = ds.attrs.find(name);
    if (it == ds.attrs.end()) fail(dataset + ": no attribute " + name);
    if (!numeric(it->second.type) || it->second.count < 1) fail(dataset + "/" + name + ": not numeric");
    return element(it->second.type, it->second.data.data());
}

int64_t H5LiteFile::attr_int(const std::string& dataset, const std::string& name) const
{
    const Dataset& ds = get(dataset);
    auto it = ds.attrs.find(name);
    if (it == ds.attrs.end()) fail(dataset + ": no attribute " + name);
    if (!numeric(it->second.type) || it->second.count < 1) fail(dataset + "/" + name + ": not numeric");
    return element_int(it->second.type, it->second.data.data());
}
