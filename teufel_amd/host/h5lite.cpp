// h5lite.cpp - minimal HDF5 reader (format subset described in h5lite.h),
// written from the HDF5 file format specification: superblock, local heap,
// version 1 B-tree (group nodes) and symbol table nodes, version 1 object
// headers with dataspace (0x0001), datatype (0x0003), data layout (0x0008),
// attribute (0x000C), continuation (0x0010) and symbol table (0x0011)
// messages.
#include "h5lite.h"

#include <algorithm>
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
            if (nodes.size() > 100000) fail("group B-tree too large or cyclic");
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
                bool overflow = false;
                a.count = 1;
                for (uint64_t x : adims) {
                    if (x != 0 && a.count > ~0ull / x) overflow = true;
                    a.count *= x;
                }
                if (overflow || (a.type.size != 0 && a.count > ~0ull / a.type.size))
                    fail("attribute " + name + ": dataspace too large");
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
    if (!numeric(it->second.type) || it->second.count < 1 || it->second.data.size() < it->second.type.size)
        fail(dataset + "/" + name + ": not numeric");
    return element(it->second.type, it->second.data.data());
}

int64_t H5LiteFile::attr_int(const std::string& dataset, const std::string& name) const
{
    const Dataset& ds = get(dataset);
    auto it = ds.attrs.find(name);
    if (it == ds.attrs.end()) fail(dataset + ": no attribute " + name);
    if (!numeric(it->second.type) || it->second.count < 1 || it->second.data.size() < it->second.type.size)
        fail(dataset + "/" + name + ": not numeric");
    return element_int(it->second.type, it->second.data.data());
}

// ---- writer -------------------------------------------------------------------

namespace {
struct Out {
    std::vector<uint8_t> b;
    void at(uint64_t off, uint64_t v, int bytes)
    {
        if (b.size() < off + bytes) b.resize(off + bytes, 0);
        for (int i = 0; i < bytes; ++i) b[off + i] = (uint8_t)(v >> (8 * i));
    }
    void raw(uint64_t off, const void* p, size_t n)
    {
        if (b.size() < off + n) b.resize(off + n, 0);
        std::memcpy(&b[off], p, n);
    }
};

// datatype message bodies: IEEE binary64 and signed 32-bit, little endian
const uint8_t kTypeF64[24] = {0x11, 0x20, 0x3f, 0x00, 8, 0, 0, 0, 0, 0, 64, 0, 52, 11, 0, 52, 0xff, 0x03, 0, 0, 0, 0, 0, 0};
const uint8_t kTypeI32[16] = {0x10, 0x08, 0x00, 0x00, 4, 0, 0, 0, 0, 0, 32, 0, 0, 0, 0, 0};
constexpr uint64_t kLeafK = 4, kInternalK = 16;
}  // namespace

H5LiteWriter::Set& H5LiteWriter::find(const std::string& name)
{
    for (Set& s : sets)
        if (s.name == name) return s;
    throw IOexception("h5lite: no dataset " + name + " to attach an attribute to");
}

void H5LiteWriter::add_double(const std::string& name, const std::vector<uint64_t>& dims, const std::vector<double>& data)
{
    uint64_t n = 1;
    for (uint64_t d : dims) n *= d;
    if (n != data.size()) throw IOexception("h5lite: " + name + ": data size does not match its dimensions");
    Set s{name, dims, true, std::vector<uint8_t>(data.size() * 8), {}};
    if (!data.empty()) std::memcpy(s.bytes.data(), data.data(), s.bytes.size());
    sets.push_back(s);
}

void H5LiteWriter::add_int32(const std::string& name, const std::vector<uint64_t>& dims, const std::vector<int32_t>& data)
{
    uint64_t n = 1;
    for (uint64_t d : dims) n *= d;
    if (n != data.size()) throw IOexception("h5lite: " + name + ": data size does not match its dimensions");
    Set s{name, dims, false, std::vector<uint8_t>(data.size() * 4), {}};
    if (!data.empty()) std::memcpy(s.bytes.data(), data.data(), s.bytes.size());
    sets.push_back(s);
}

void H5LiteWriter::attr_int32(const std::string& dataset, const std::string& name, int32_t value)
{
    find(dataset).attrs.push_back(Attr{name, false, 0.0, value});
}

void H5LiteWriter::attr_double(const std::string& dataset, const std::string& name, double value)
{
    find(dataset).attrs.push_back(Attr{name, true, value, 0});
}

void H5LiteWriter::write(const std::string& filename) const
{
    if (sets.size() > 2 * kLeafK) throw IOexception("h5lite: at most 8 datasets per file");
    std::vector<const Set*> order;
    for (const Set& s : sets) order.push_back(&s);
    std::sort(order.begin(), order.end(), [](const Set* a, const Set* b) { return a->name < b->name; });
    for (size_t i = 1; i < order.size(); ++i)
        if (order[i]->name == order[i - 1]->name) throw IOexception("h5lite: duplicate dataset " + order[i]->name);

    // local heap data: "" at 0, then the names, each padded to 8 bytes
    std::vector<uint8_t> heap(8, 0);
    std::vector<uint64_t> name_off;
    for (const Set* s : order) {
        name_off.push_back(heap.size());
        heap.insert(heap.end(), s->name.begin(), s->name.end());
        heap.resize(pad8(heap.size() + 1), 0);
    }
    // object header message bodies of each dataset
    auto dataspace = [](const std::vector<uint64_t>& dims) {
        std::vector<uint8_t> m(8 + 8 * dims.size(), 0);
        m[0] = 1;
        m[1] = (uint8_t)dims.size();
        for (size_t i = 0; i < dims.size(); ++i)
            for (int k = 0; k < 8; ++k) m[8 + 8 * i + k] = (uint8_t)(dims[i] >> (8 * k));
        return m;
    };
    struct Msg {
        uint16_t type;
        uint8_t flags;
        std::vector<uint8_t> body;
    };
    std::vector<std::vector<Msg>> msgs(order.size());
    for (size_t i = 0; i < order.size(); ++i) {
        const Set* s = order[i];
        msgs[i].push_back({0x0001, 0, dataspace(s->dims)});
        msgs[i].push_back(s->is_double ? Msg{0x0003, 1, std::vector<uint8_t>(kTypeF64, kTypeF64 + 20)}
                                       : Msg{0x0003, 1, std::vector<uint8_t>(kTypeI32, kTypeI32 + 12)});
        msgs[i].push_back({0x0008, 0, std::vector<uint8_t>(18, 0)});  // layout, addresses patched below
        for (const Attr& a : s->attrs) {
            std::vector<uint8_t> m(8, 0);
            const size_t nlen = a.name.size() + 1;
            const size_t tlen = a.is_double ? 20 : 12;
            m[0] = 1;
            m[2] = (uint8_t)nlen;
            m[3] = (uint8_t)(nlen >> 8);
            m[4] = (uint8_t)tlen;
            m[6] = 8;  // scalar dataspace message
            std::vector<uint8_t> name(pad8(nlen), 0);
            std::memcpy(name.data(), a.name.c_str(), a.name.size());
            m.insert(m.end(), name.begin(), name.end());
            std::vector<uint8_t> t(a.is_double ? kTypeF64 : kTypeI32, (a.is_double ? kTypeF64 : kTypeI32) + tlen);
            t.resize(pad8(tlen), 0);
            m.insert(m.end(), t.begin(), t.end());
            std::vector<uint8_t> sp(8, 0);
            sp[0] = 1;
            m.insert(m.end(), sp.begin(), sp.end());
            uint8_t v[8] = {0};
            if (a.is_double)
                std::memcpy(v, &a.d, 8);
            else
                std::memcpy(v, &a.i, 4);
            m.insert(m.end(), v, v + (a.is_double ? 8 : 4));
            msgs[i].push_back({0x000C, 0, m});
        }
    }
    auto header_size = [&](const std::vector<Msg>& ms) {
        uint64_t n = 0;
        for (const Msg& m : ms) n += 8 + pad8(m.body.size());
        return n;
    };
    // addresses
    const uint64_t root_oh = 96;
    const uint64_t root_oh_size = 16 + 8 + 16;
    const uint64_t heap_hdr = pad8(root_oh + root_oh_size);
    const uint64_t heap_data = heap_hdr + 32;
    const uint64_t btree = pad8(heap_data + heap.size());
    const uint64_t btree_size = 24 + (2 * kInternalK + 1) * 8 + 2 * kInternalK * 8;
    const uint64_t snod = pad8(btree + btree_size);
    const uint64_t snod_size = 8 + 2 * kLeafK * 40;
    std::vector<uint64_t> oh(order.size()), data(order.size());
    uint64_t p = pad8(snod + snod_size);
    for (size_t i = 0; i < order.size(); ++i) {
        oh[i] = p;
        p = pad8(p + 16 + header_size(msgs[i]));
    }
    for (size_t i = 0; i < order.size(); ++i) {
        data[i] = p;
        p = pad8(p + order[i]->bytes.size());
    }
    const uint64_t eof = p;

    Out o;
    o.raw(0, kSignature, 8);
    o.at(13, 8, 1);
    o.at(14, 8, 1);
    o.at(16, kLeafK, 2);
    o.at(18, kInternalK, 2);
    o.at(24, 0, 8);
    o.at(32, kUndefined, 8);
    o.at(40, eof, 8);
    o.at(48, kUndefined, 8);
    o.at(56, 0, 8);          // root entry: link name offset
    o.at(64, root_oh, 8);
    o.at(72, 1, 4);          // cache type 1: B-tree and heap in the scratch pad
    o.at(80, btree, 8);
    o.at(88, heap_hdr, 8);
    // root object header: one symbol table message
    o.at(root_oh, 1, 1);
    o.at(root_oh + 2, 1, 2);
    o.at(root_oh + 4, 1, 4);
    o.at(root_oh + 8, 8 + 16, 4);
    o.at(root_oh + 16, 0x0011, 2);
    o.at(root_oh + 18, 16, 2);
    o.at(root_oh + 24, btree, 8);
    o.at(root_oh + 32, heap_hdr, 8);
    // local heap
    o.raw(heap_hdr, "HEAP", 4);
    o.at(heap_hdr + 8, heap.size(), 8);
    o.at(heap_hdr + 16, 1, 8);  // free-list head: 1 is the "no free block" marker
    o.at(heap_hdr + 24, heap_data, 8);
    o.raw(heap_data, heap.data(), heap.size());
    // group B-tree: one child, the symbol node
    o.raw(btree, "TREE", 4);
    o.at(btree + 6, order.empty() ? 0 : 1, 2);
    o.at(btree + 8, kUndefined, 8);
    o.at(btree + 16, kUndefined, 8);
    o.at(btree + 24, 0, 8);
    o.at(btree + 32, snod, 8);
    o.at(btree + 40, order.empty() ? 0 : name_off.back(), 8);
    o.at(btree + btree_size - 1, 0, 1);
    // symbol table node
    o.raw(snod, "SNOD", 4);
    o.at(snod + 4, 1, 1);
    o.at(snod + 6, order.size(), 2);
    for (size_t i = 0; i < order.size(); ++i) {
        o.at(snod + 8 + 40 * i, name_off[i], 8);
        o.at(snod + 16 + 40 * i, oh[i], 8);
    }
    o.at(snod + snod_size - 1, 0, 1);
    // dataset object headers and data
    for (size_t i = 0; i < order.size(); ++i) {
        std::vector<Msg>& ms = msgs[i];
        std::vector<uint8_t>& lay = ms[2].body;
        lay[0] = 3;
        lay[1] = 1;
        for (int k = 0; k < 8; ++k) lay[2 + k] = (uint8_t)(data[i] >> (8 * k));
        for (int k = 0; k < 8; ++k) lay[10 + k] = (uint8_t)((uint64_t)order[i]->bytes.size() >> (8 * k));
        o.at(oh[i], 1, 1);
        o.at(oh[i] + 2, ms.size(), 2);
        o.at(oh[i] + 4, 1, 4);
        o.at(oh[i] + 8, header_size(ms), 4);
        uint64_t q = oh[i] + 16;
        for (const Msg& m : ms) {
            o.at(q, m.type, 2);
            o.at(q + 2, pad8(m.body.size()), 2);
            o.at(q + 4, m.flags, 1);
            o.raw(q + 8, m.body.data(), m.body.size());
            q += 8 + pad8(m.body.size());
        }
        if (!order[i]->bytes.empty()) o.raw(data[i], order[i]->bytes.data(), order[i]->bytes.size());
    }
    o.at(eof - 1, o.b.size() >= eof ? o.b[eof - 1] : 0, 1);

    FILE* f = std::fopen(filename.c_str(), "wb");
    if (!f) throw IOexception("h5lite: cannot create " + filename);
    const bool ok = std::fwrite(o.b.data(), 1, o.b.size(), f) == o.b.size();
    if (std::fclose(f) != 0 || !ok) throw IOexception("h5lite: write error on " + filename);
}
