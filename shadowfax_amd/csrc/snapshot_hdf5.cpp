// snapshot_hdf5.cpp -- a reader for the HDF5 files of Gadget-style snapshots
// and initial conditions (the files GadgetSnapshotReader reads through the
// HDF5 library, src/GadgetSnapshotReader.cpp:79-380), written from the
// published HDF5 file format specification for the structures such files
// use: superblock version 0, groups as symbol tables (version 1 B-tree,
// local heap, symbol table nodes), version 1 object headers with their
// continuation blocks, dataspace, datatype, data layout (compact and
// contiguous) and attribute messages.  Chunked or filtered datasets, shared
// messages and version 2 object headers are reported as unsupported.
// Numbers are converted to double or int64 on reading; integer and floating
// types of 1 to 8 bytes, little-endian, and enumerations over them (booleans)
// are understood.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "sfx_capi_internal.h"
#include "shadowfax_amd.h"

using sfx::set_error;

namespace {

struct Type {
    int cls = -1;  // 0 fixed point, 1 floating point, 8 enum (of a fixed-point base)
    int size = 0;
    bool is_signed = false;
};

struct Space {
    int rank = 0;
    bool empty = false;  // a null dataspace: no elements
    std::vector<uint64_t> dims;
    uint64_t count() const {
        if(empty) return 0;
        uint64_t n = 1;
        for(uint64_t d : dims) n *= d;
        return n;
    }
};

struct Attr {
    Type type;
    Space space;
    uint64_t data = 0;  // file offset of the values
};

struct Object {
    bool is_group = false, is_dataset = false;
    uint64_t btree = 0, heap = 0;  // group
    Type type;                     // dataset
    Space space;
    int layout = -1;  // 0 compact, 1 contiguous
    uint64_t data = 0, data_size = 0;
    std::map<std::string, Attr> attrs;
};

struct Error {
    std::string what;
};

}  // namespace

struct sfx_h5 {
    std::vector<uint8_t> buf;
    int O = 8, L = 8;  // sizes of offsets and lengths
    uint64_t base = 0;
    uint64_t root = 0;  // object header address of the root group

    void need(uint64_t off, uint64_t n) const {
        if(off > buf.size() || n > buf.size() - off) throw Error{"read past the end of the file"};
    }
    uint64_t u(uint64_t off, int n) const {
        need(off, (uint64_t)n);
        uint64_t v = 0;
        for(int k = n; k--;) v = (v << 8) | buf[off + k];
        return v;
    }
    uint64_t addr(uint64_t off) const { return u(off, O); }
    uint64_t abs_(uint64_t a) const { return base + a; }
    bool undefined(uint64_t a) const { return O == 8 ? a == ~uint64_t(0) : a == (uint64_t(1) << (8 * O)) - 1; }

    // a datatype message at off; returns its length in bytes
    uint64_t parse_type(uint64_t off, Type& t) const {
        const int cls = (int)(u(off, 1) & 15);
        const uint64_t bits = u(off + 1, 3);
        t.size = (int)u(off + 4, 4);
        if(cls == 0) {
            if(bits & 1) throw Error{"big-endian integers are not supported"};
            if(t.size < 1 || t.size > 8) throw Error{"integers of 1 to 8 bytes only"};
            t.cls = 0;
            t.is_signed = (bits >> 3) & 1;
            return 8 + 4;
        }
        if(cls == 1) {
            if((bits & 1) || ((bits >> 6) & 1)) throw Error{"big-endian floats are not supported"};
            if(t.size != 4 && t.size != 8) throw Error{"floats of 4 or 8 bytes only"};
            t.cls = 1;
            return 8 + 12;
        }
        if(cls == 8) {
            Type b;
            parse_type(off + 8, b);
            if(b.cls != 0) throw Error{"enumerations over integers only"};
            t = b;
            t.cls = 8;
            return 0;  // not needed: enums only appear as whole messages
        }
        throw Error{"datatype class " + std::to_string(cls) + " is not supported"};
    }

    void parse_space(uint64_t off, Space& s) const {
        const int ver = (int)u(off, 1);
        s.rank = (int)u(off + 1, 1);
        uint64_t p;
        if(ver == 1) {
            p = off + 8;
        } else if(ver == 2) {
            if(u(off + 3, 1) == 2) {
                s.rank = 0;
                s.empty = true;
            }
            p = off + 4;
        } else {
            throw Error{"dataspace version " + std::to_string(ver)};
        }
        s.dims.assign((size_t)s.rank, 0);
        for(int d = 0; d < s.rank; d++) s.dims[d] = u(p + (uint64_t)d * L, L);
    }

    // the messages of a version 1 object header and its continuation blocks
    Object parse_object(uint64_t a) const {
        const uint64_t off = abs_(a);
        if(u(off, 1) != 1) throw Error{"object header version other than 1"};
        const uint64_t nmsg = u(off + 2, 2);
        const uint64_t hsize = u(off + 8, 4);
        std::vector<std::pair<uint64_t, uint64_t>> blocks{{off + 16, hsize}};
        Object o;
        uint64_t seen = 0;
        for(size_t b = 0; b < blocks.size() && seen < nmsg; b++) {
            uint64_t p = blocks[b].first;
            const uint64_t end = blocks[b].first + blocks[b].second;
            while(p + 8 <= end && seen < nmsg) {
                const int type = (int)u(p, 2);
                const uint64_t size = u(p + 2, 2);
                const int flags = (int)u(p + 4, 1);
                const uint64_t d = p + 8;
                need(d, size);
                seen++;
                if((flags & 2) && type != 0) throw Error{"shared header messages are not supported"};
                switch(type) {
                    case 0x0001: parse_space(d, o.space); break;
                    case 0x0003: parse_type(d, o.type); break;
                    case 0x0008: parse_layout(d, o); break;
                    case 0x000B: throw Error{"filtered (compressed) datasets are not supported"};
                    case 0x000C: parse_attr(d, o); break;
                    case 0x0010:
                        blocks.push_back({abs_(addr(d)), u(d + O, L)});
                        break;
                    case 0x0011:
                        o.is_group = true;
                        o.btree = addr(d);
                        o.heap = addr(d + O);
                        break;
                    default: break;  // fill value, times, NIL, ...
                }
                p = d + size;
            }
        }
        o.is_dataset = !o.is_group && o.layout >= 0;
        return o;
    }

    void parse_layout(uint64_t d, Object& o) const {
        const int ver = (int)u(d, 1);
        if(ver == 3 || ver == 4) {
            const int cls = (int)u(d + 1, 1);
            if(cls == 0) {
                o.layout = 0;
                o.data_size = u(d + 2, 2);
                o.data = d + 4;
            } else if(cls == 1) {
                o.layout = 1;
                o.data = addr(d + 2);
                o.data_size = u(d + 2 + O, L);
            } else {
                throw Error{"chunked datasets are not supported"};
            }
        } else if(ver == 1 || ver == 2) {
            const int nd = (int)u(d + 1, 1), cls = (int)u(d + 2, 1);
            if(cls == 0) {
                o.layout = 0;
                const uint64_t q = d + 8 + 4 * (uint64_t)nd;
                o.data_size = u(q, 4);
                o.data = q + 4;
            } else if(cls == 1) {
                o.layout = 1;
                o.data = addr(d + 8);
                uint64_t n = 1;
                for(int k = 0; k < nd; k++) n *= u(d + 8 + O + 4 * (uint64_t)k, 4);
                o.data_size = n;  // elements x element size (the last dimension)
            } else {
                throw Error{"chunked datasets are not supported"};
            }
        } else {
            throw Error{"data layout version " + std::to_string(ver)};
        }
    }

    void parse_attr(uint64_t d, Object& o) const {
        const int ver = (int)u(d, 1);
        const uint64_t nname = u(d + 2, 2), ntype = u(d + 4, 2), nspace = u(d + 6, 2);
        auto pad8 = [](uint64_t x) { return (x + 7) & ~uint64_t(7); };
        uint64_t p;
        uint64_t name_len = nname, type_len = ntype, space_len = nspace;
        if(ver == 1) {
            p = d + 8;
            name_len = pad8(nname);
            type_len = pad8(ntype);
            space_len = pad8(nspace);
        } else if(ver == 2) {
            p = d + 8;
        } else if(ver == 3) {
            p = d + 9;
        } else {
            throw Error{"attribute message version " + std::to_string(ver)};
        }
        need(p, nname);
        std::string name((const char*)&buf[p], nname ? nname - 1 : 0);
        Attr a;
        try {
            parse_type(p + name_len, a.type);
        } catch(const Error&) {
            a.type.cls = -1;  // strings and other types: listed, not readable
        }
        parse_space(p + name_len + type_len, a.space);
        a.data = p + name_len + type_len + space_len;
        o.attrs[name] = a;
    }

    std::string heap_name(uint64_t heap, uint64_t off) const {
        const uint64_t h = abs_(heap);
        need(h, 8 + 3 * (uint64_t)L);  // signature, sizes and the segment address
        if(std::memcmp(&buf[h], "HEAP", 4) != 0) throw Error{"bad local heap"};
        const uint64_t seg = abs_(addr(h + 8 + 2 * (uint64_t)L));
        uint64_t p = seg + off;
        std::string s;
        while(true) {
            need(p, 1);
            if(!buf[p]) break;
            s.push_back((char)buf[p++]);
        }
        return s;
    }

    void walk_btree(uint64_t node, uint64_t heap, std::map<std::string, uint64_t>& out,
                    int depth) const {
        if(depth > 32) throw Error{"group B-tree too deep"};
        const uint64_t n = abs_(node);
        need(n, 8);
        if(std::memcmp(&buf[n], "TREE", 4) != 0) throw Error{"bad group B-tree node"};
        const int level = (int)u(n + 5, 1);
        const uint64_t used = u(n + 6, 2);
        const uint64_t first = n + 8 + 2 * (uint64_t)O;
        for(uint64_t k = 0; k < used; k++) {
            const uint64_t child = addr(first + (uint64_t)L + k * (uint64_t)(L + O));
            if(level > 0) {
                walk_btree(child, heap, out, depth + 1);
                continue;
            }
            const uint64_t s = abs_(child);
            need(s, 8);
            if(std::memcmp(&buf[s], "SNOD", 4) != 0) throw Error{"bad symbol table node"};
            const uint64_t nsym = u(s + 6, 2);
            const uint64_t esize = 2 * (uint64_t)O + 24;
            for(uint64_t e = 0; e < nsym; e++) {
                const uint64_t ent = s + 8 + e * esize;
                out[heap_name(heap, addr(ent))] = addr(ent + O);
            }
        }
    }

    // object header address of the object at an absolute path ("/A/B")
    bool lookup(const std::string& path, uint64_t& a) const {
        a = root;
        size_t i = 0;
        while(i < path.size()) {
            while(i < path.size() && path[i] == '/') i++;
            if(i >= path.size()) break;
            size_t j = path.find('/', i);
            if(j == std::string::npos) j = path.size();
            const std::string part = path.substr(i, j - i);
            const Object g = parse_object(a);
            if(!g.is_group) return false;
            std::map<std::string, uint64_t> links;
            walk_btree(g.btree, g.heap, links, 0);
            auto it = links.find(part);
            if(it == links.end()) return false;
            a = it->second;
            i = j;
        }
        return true;
    }

    // n values of type t at absolute offset p, converted
    template <typename Out>
    void convert(uint64_t p, const Type& t, uint64_t n, Out* out) const {
        need(p, n * (uint64_t)t.size);
        for(uint64_t k = 0; k < n; k++) {
            const uint64_t q = p + k * (uint64_t)t.size;
            if(t.cls == 1) {
                if(t.size == 8) {
                    double v;
                    std::memcpy(&v, &buf[q], 8);
                    out[k] = (Out)v;
                } else {
                    float v;
                    std::memcpy(&v, &buf[q], 4);
                    out[k] = (Out)v;
                }
            } else {
                uint64_t v = u(q, t.size);
                if(t.is_signed && t.size < 8 && ((v >> (8 * t.size - 1)) & 1))
                    v |= ~uint64_t(0) << (8 * t.size);
                out[k] = t.is_signed ? (Out)(int64_t)v : (Out)v;
            }
        }
    }
};

namespace {

int h5_fail(const char* who, const Error& e) {
    return set_error(SFX_ERR_ARG, "%s: %s", who, e.what.c_str());
}

// the object at path and, if name is given, its attribute
int find(const sfx_h5* f, const char* who, const char* path, const char* name, Object& o,
         const Attr** a) {
    uint64_t at = 0;
    if(!f->lookup(path ? path : "/", at))
        return set_error(SFX_ERR_ARG, "%s: no object %s", who, path ? path : "/");
    o = f->parse_object(at);
    if(name) {
        auto it = o.attrs.find(name);
        if(it == o.attrs.end())
            return set_error(SFX_ERR_ARG, "%s: %s has no attribute %s", who, path, name);
        *a = &it->second;
    }
    return SFX_OK;
}

}  // namespace

extern "C" {

int sfx_h5_open(const char* path, sfx_h5** out) {
    if(!path || !out) return set_error(SFX_ERR_ARG, "sfx_h5_open: NULL");
    *out = nullptr;
    FILE* fp = std::fopen(path, "rb");
    if(!fp) return set_error(SFX_ERR_ARG, "sfx_h5_open: cannot open %s", path);
    sfx_h5* f = new sfx_h5();
    std::fseek(fp, 0, SEEK_END);
    const long size = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    f->buf.resize(size > 0 ? (size_t)size : 0);
    const size_t got = size > 0 ? std::fread(f->buf.data(), 1, (size_t)size, fp) : 0;
    std::fclose(fp);
    if(got != f->buf.size()) {
        delete f;
        return set_error(SFX_ERR_ARG, "sfx_h5_open: short read of %s", path);
    }
    static const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    uint64_t sb = 0;
    bool found = false;
    for(uint64_t at = 0; at + 8 <= f->buf.size(); at = at ? 2 * at : 512) {
        if(std::memcmp(&f->buf[at], sig, 8) == 0) {
            sb = at;
            found = true;
            break;
        }
    }
    try {
        if(!found) throw Error{"no HDF5 signature"};
        if(f->u(sb + 8, 1) != 0) throw Error{"superblock version other than 0"};
        f->O = (int)f->u(sb + 13, 1);
        f->L = (int)f->u(sb + 14, 1);
        if(f->O < 2 || f->O > 8 || f->L < 2 || f->L > 8) throw Error{"bad offset sizes"};
        const uint64_t p = sb + 24;  // base address, free space, end of file, driver info
        f->base = f->addr(p);
        const uint64_t root_entry = p + 4 * (uint64_t)f->O;
        f->root = f->addr(root_entry + f->O);
        (void)f->parse_object(f->root);
    } catch(const Error& e) {
        delete f;
        return h5_fail("sfx_h5_open", e);
    }
    *out = f;
    return SFX_OK;
}

int sfx_h5_close(sfx_h5* f) {
    delete f;
    return SFX_OK;
}

int sfx_h5_exists(const sfx_h5* f, const char* path, const char* attr) {
    if(!f || !path) return set_error(SFX_ERR_ARG, "sfx_h5_exists: NULL");
    try {
        uint64_t a = 0;
        if(!f->lookup(path, a)) return 0;
        if(!attr) return 1;
        const Object o = f->parse_object(a);
        return o.attrs.count(attr) ? 1 : 0;
    } catch(const Error& e) {
        return h5_fail("sfx_h5_exists", e);
    }
}

int sfx_h5_info(const sfx_h5* f, const char* path, const char* attr, int* type_class,
                int* type_size, int* rank, int64_t* dims) {
    if(!f || !path) return set_error(SFX_ERR_ARG, "sfx_h5_info: NULL");
    try {
        Object o;
        const Attr* a = nullptr;
        int rc = find(f, "sfx_h5_info", path, attr, o, &a);
        if(rc) return rc;
        const Type& t = a ? a->type : o.type;
        const Space& s = a ? a->space : o.space;
        if(!a && !o.is_dataset) return set_error(SFX_ERR_ARG, "sfx_h5_info: %s is a group", path);
        if(type_class) *type_class = t.cls;
        if(type_size) *type_size = t.size;
        if(rank) *rank = s.rank;
        if(dims)
            for(int d = 0; d < s.rank && d < 8; d++) dims[d] = (int64_t)s.dims[d];
        return SFX_OK;
    } catch(const Error& e) {
        return h5_fail("sfx_h5_info", e);
    }
}

static int h5_read(const sfx_h5* f, const char* path, const char* attr, int64_t n, double* d,
                   int64_t* i) {
    const char* who = "sfx_h5_read";
    if(!f || !path || (!d && !i) || n < 0) return set_error(SFX_ERR_ARG, "%s: bad arguments", who);
    try {
        Object o;
        const Attr* a = nullptr;
        int rc = find(f, who, path, attr, o, &a);
        if(rc) return rc;
        if(!a && !o.is_dataset) return set_error(SFX_ERR_ARG, "%s: %s is a group", who, path);
        const Type& t = a ? a->type : o.type;
        const Space& s = a ? a->space : o.space;
        if(t.cls < 0) return set_error(SFX_ERR_ARG, "%s: %s: type not readable as numbers", who, path);
        const uint64_t count = s.count();
        if((uint64_t)n != count)
            return set_error(SFX_ERR_ARG, "%s: %s holds %llu values, %lld asked", who, path,
                             (unsigned long long)count, (long long)n);
        uint64_t p;
        if(a) {
            p = a->data;
        } else {
            if(o.data_size < count * (uint64_t)t.size && o.layout == 0)
                return set_error(SFX_ERR_ARG, "%s: %s: compact data too short", who, path);
            p = o.layout == 0 ? o.data : f->abs_(o.data);
            if(o.layout == 1 && f->undefined(o.data) && count)
                return set_error(SFX_ERR_ARG, "%s: %s has no storage", who, path);
        }
        if(d)
            f->convert(p, t, count, d);
        else
            f->convert(p, t, count, i);
        return SFX_OK;
    } catch(const Error& e) {
        return h5_fail(who, e);
    }
}

int sfx_h5_read_f64(const sfx_h5* f, const char* path, const char* attr, int64_t n, double* out) {
    return h5_read(f, path, attr, n, out, nullptr);
}

int sfx_h5_read_i64(const sfx_h5* f, const char* path, const char* attr, int64_t n,
                    int64_t* out) {
    return h5_read(f, path, attr, n, nullptr, out);
}

int sfx_h5_list(const sfx_h5* f, const char* path, char* names, int64_t cap) {
    if(!f || !path || (!names && cap)) return set_error(SFX_ERR_ARG, "sfx_h5_list: NULL");
    try {
        Object o;
        const Attr* a = nullptr;
        int rc = find(f, "sfx_h5_list", path, nullptr, o, &a);
        if(rc) return rc;
        std::string all;
        if(o.is_group) {
            std::map<std::string, uint64_t> links;
            f->walk_btree(o.btree, o.heap, links, 0);
            for(auto& kv : links) all += kv.first + "\n";
        }
        for(auto& kv : o.attrs) all += "@" + kv.first + "\n";
        if((int64_t)all.size() + 1 > cap) return (int)(all.size() + 1);
        std::memcpy(names, all.c_str(), all.size() + 1);
        return SFX_OK;
    } catch(const Error& e) {
        return h5_fail("sfx_h5_list", e);
    }
}

}  // extern "C"

// ------------------------------------------------------------ writing --
//
// The same structures, written: superblock version 0 (offsets and lengths of
// 8 bytes, group leaf node K 4, internal node K 16), each group a version 1
// object header holding a symbol table message and its attributes, a
// version 1 B-tree leaf of full size (room for 2K = 32 children) whose
// children are symbol table nodes of up to 2 x 4 entries each (names sorted
// across them; key j + 1 is the last name of child j, key 0 the empty
// name), a local heap whose data segment starts with the empty name; each
// dataset a version 1 object header with dataspace, datatype, contiguous
// layout and attribute messages, its data 8-byte aligned after it.  At most
// 256 members per group.

namespace {

struct WNode {
    std::string name;
    bool group = true;
    int cls = 1, size = 8;
    std::vector<uint64_t> dims;
    std::vector<uint8_t> data;
    struct WAttr {
        std::string name;
        int cls, size;
        std::vector<uint64_t> dims;
        std::vector<uint8_t> data;
    };
    std::vector<WAttr> attrs;
    std::vector<WNode*> kids;
    // layout
    uint64_t oh = 0, oh_size = 0, btree = 0, heap = 0, heap_data = 0, heap_size = 0, raw = 0;
    std::vector<uint64_t> snods;  // one per 8 members
    std::vector<uint64_t> name_off;  // heap offsets of the (sorted) kids' names
    ~WNode() {
        for(WNode* k : kids) delete k;
    }
};

constexpr int kLeafK = 4, kInternalK = 16;
constexpr size_t kSnodEntries = 2 * kLeafK, kMaxMembers = 2 * kInternalK * kSnodEntries;
constexpr uint64_t kUndef = ~uint64_t(0);

uint64_t pad8(uint64_t x) { return (x + 7) & ~uint64_t(7); }

struct Out {
    std::vector<uint8_t> b;
    void at(uint64_t off, uint64_t v, int n) {
        if(b.size() < off + n) b.resize(off + n, 0);
        for(int k = 0; k < n; k++) b[off + k] = (uint8_t)(v >> (8 * k));
    }
    void bytes(uint64_t off, const void* p, uint64_t n) {
        if(b.size() < off + n) b.resize(off + n, 0);
        if(n) std::memcpy(&b[off], p, n);
    }
};

// a datatype message (unpadded bytes)
std::vector<uint8_t> type_msg(int cls, int size) {
    std::vector<uint8_t> m(cls == 1 ? 20 : 12, 0);
    auto put = [&](int off, uint64_t v, int n) {
        for(int k = 0; k < n; k++) m[off + k] = (uint8_t)(v >> (8 * k));
    };
    if(cls == 1) {
        m[0] = 0x11;  // version 1, floating point
        m[1] = 0x20;  // little-endian, implied mantissa bit
        m[2] = (uint8_t)(8 * size - 1);  // sign bit
        put(4, (uint64_t)size, 4);
        put(8, 0, 2);
        put(10, (uint64_t)(8 * size), 2);
        if(size == 8) {
            m[12] = 52, m[13] = 11, m[14] = 0, m[15] = 52;
            put(16, 1023, 4);
        } else {
            m[12] = 23, m[13] = 8, m[14] = 0, m[15] = 23;
            put(16, 127, 4);
        }
    } else {
        m[0] = 0x10;                  // version 1, fixed point
        m[1] = cls == 2 ? 0x08 : 0;   // cls 2 here: signed integer
        put(4, (uint64_t)size, 4);
        put(8, 0, 2);
        put(10, (uint64_t)(8 * size), 2);
    }
    return m;
}

std::vector<uint8_t> space_msg(const std::vector<uint64_t>& dims) {
    std::vector<uint8_t> m(8 + 8 * dims.size(), 0);
    m[0] = 1;
    m[1] = (uint8_t)dims.size();
    for(size_t d = 0; d < dims.size(); d++)
        for(int k = 0; k < 8; k++) m[8 + 8 * d + k] = (uint8_t)(dims[d] >> (8 * k));
    return m;
}

std::vector<uint8_t> attr_msg(const WNode::WAttr& a) {
    const std::vector<uint8_t> t = type_msg(a.cls, a.size), s = space_msg(a.dims);
    const uint64_t nn = a.name.size() + 1;
    std::vector<uint8_t> m(8 + pad8(nn) + pad8(t.size()) + pad8(s.size()) + a.data.size(), 0);
    m[0] = 1;
    m[2] = (uint8_t)nn, m[3] = (uint8_t)(nn >> 8);
    m[4] = (uint8_t)t.size(), m[5] = (uint8_t)(t.size() >> 8);
    m[6] = (uint8_t)s.size(), m[7] = (uint8_t)(s.size() >> 8);
    uint64_t p = 8;
    std::memcpy(&m[p], a.name.c_str(), a.name.size());
    p += pad8(nn);
    std::memcpy(&m[p], t.data(), t.size());
    p += pad8(t.size());
    std::memcpy(&m[p], s.data(), s.size());
    p += pad8(s.size());
    if(!a.data.empty()) std::memcpy(&m[p], a.data.data(), a.data.size());
    return m;
}

// the messages of an object header as (type, bytes)
std::vector<std::pair<int, std::vector<uint8_t>>> messages(const WNode* n) {
    std::vector<std::pair<int, std::vector<uint8_t>>> ms;
    if(n->group) {
        std::vector<uint8_t> st(16, 0);
        for(int k = 0; k < 8; k++) {
            st[k] = (uint8_t)(n->btree >> (8 * k));
            st[8 + k] = (uint8_t)(n->heap >> (8 * k));
        }
        ms.push_back({0x0011, st});
    } else {
        ms.push_back({0x0001, space_msg(n->dims)});
        ms.push_back({0x0003, type_msg(n->cls, n->size)});
        std::vector<uint8_t> lay(18, 0);
        lay[0] = 3;
        lay[1] = 1;  // contiguous
        const uint64_t a = n->data.empty() ? kUndef : n->raw;
        for(int k = 0; k < 8; k++) {
            lay[2 + k] = (uint8_t)(a >> (8 * k));
            lay[10 + k] = (uint8_t)((uint64_t)n->data.size() >> (8 * k));
        }
        ms.push_back({0x0008, lay});
    }
    for(const auto& a : n->attrs) ms.push_back({0x000C, attr_msg(a)});
    return ms;
}

uint64_t oh_bytes(const WNode* n) {
    uint64_t s = 16;
    for(auto& m : messages(n)) s += 8 + pad8(m.second.size());
    return s;
}

constexpr uint64_t kBtreeBytes = 8 + 2 * 8 + 2 * kInternalK * 8 + (2 * kInternalK + 1) * 8;
constexpr uint64_t kSnodBytes = 8 + 2 * kLeafK * 40;

// addresses, depth first
void place(WNode* n, uint64_t& at) {
    n->oh = at;
    // sizes of the messages do not depend on the addresses
    n->oh_size = oh_bytes(n);
    at += n->oh_size;
    if(n->group) {
        std::sort(n->kids.begin(), n->kids.end(),
                  [](const WNode* a, const WNode* b) { return a->name < b->name; });
        n->btree = at;
        at += kBtreeBytes;
        n->snods.clear();
        for(size_t k = 0; k < n->kids.size(); k += kSnodEntries) {
            n->snods.push_back(at);
            at += kSnodBytes;
        }
        n->heap = at;
        at += 32;
        n->heap_data = at;
        uint64_t h = 8;  // offset 0: the empty name
        n->name_off.clear();
        for(WNode* k : n->kids) {
            n->name_off.push_back(h);
            h += pad8(k->name.size() + 1);
        }
        n->heap_size = h;
        at += h;
        for(WNode* k : n->kids) place(k, at);
    } else {
        n->raw = at;
        at += pad8(n->data.size());
    }
}

void emit(const WNode* n, Out& o) {
    uint64_t p = n->oh;
    const auto ms = messages(n);
    o.at(p, 1, 1);
    o.at(p + 2, ms.size(), 2);
    o.at(p + 4, 1, 4);
    o.at(p + 8, n->oh_size - 16, 4);
    p += 16;
    for(auto& m : ms) {
        o.at(p, (uint64_t)m.first, 2);
        o.at(p + 2, pad8(m.second.size()), 2);
        o.bytes(p + 8, m.second.data(), m.second.size());
        p += 8 + pad8(m.second.size());
    }
    if(n->group) {
        // B-tree leaf: key 0 the empty name, child j a symbol table node,
        // key j + 1 the last name in child j
        o.bytes(n->btree, "TREE", 4);
        o.at(n->btree + 4, 0, 1);
        o.at(n->btree + 5, 0, 1);
        o.at(n->btree + 6, n->snods.size(), 2);
        o.at(n->btree + 8, kUndef, 8);
        o.at(n->btree + 16, kUndef, 8);
        o.at(n->btree + 24, 0, 8);
        for(size_t j = 0; j < n->snods.size(); j++) {
            const size_t last = std::min(n->kids.size(), (j + 1) * kSnodEntries) - 1;
            o.at(n->btree + 32 + 16 * j, n->snods[j], 8);
            o.at(n->btree + 40 + 16 * j, n->name_off[last], 8);
        }
        for(size_t j = 0; j < n->snods.size(); j++) {
            const uint64_t sn = n->snods[j];
            const size_t k0 = j * kSnodEntries,
                         k1 = std::min(n->kids.size(), k0 + kSnodEntries);
            o.bytes(sn, "SNOD", 4);
            o.at(sn + 4, 1, 1);
            o.at(sn + 6, k1 - k0, 2);
            for(size_t k = k0; k < k1; k++) {
                const uint64_t e = sn + 8 + 40 * (k - k0);
                const WNode* c = n->kids[k];
                o.at(e, n->name_off[k], 8);
                o.at(e + 8, c->oh, 8);
                o.at(e + 16, c->group ? 1 : 0, 4);
                if(c->group) {
                    o.at(e + 24, c->btree, 8);
                    o.at(e + 32, c->heap, 8);
                }
            }
        }
        o.bytes(n->heap, "HEAP", 4);
        o.at(n->heap + 8, n->heap_size, 8);
        o.at(n->heap + 16, 1, 8);  // no free block (the HDF5 library's value for none)
        o.at(n->heap + 24, n->heap_data, 8);
        for(size_t k = 0; k < n->kids.size(); k++)
            o.bytes(n->heap_data + n->name_off[k], n->kids[k]->name.c_str(),
                    n->kids[k]->name.size());
        for(const WNode* c : n->kids) emit(c, o);
    } else if(!n->data.empty()) {
        o.bytes(n->raw, n->data.data(), n->data.size());
    }
}

}  // namespace

struct sfx_h5w {
    WNode root;
    // the node at an absolute path; with create, missing groups are made
    WNode* find(const std::string& path, bool create_last_group, const char* who) {
        WNode* n = &root;
        size_t i = 0;
        while(i < path.size()) {
            while(i < path.size() && path[i] == '/') i++;
            if(i >= path.size()) break;
            size_t j = path.find('/', i);
            if(j == std::string::npos) j = path.size();
            const std::string part = path.substr(i, j - i);
            if(!n->group) throw Error{std::string(who) + ": a dataset has no members"};
            WNode* next = nullptr;
            for(WNode* k : n->kids)
                if(k->name == part) next = k;
            if(!next) {
                if(!(create_last_group && j == path.size()))
                    throw Error{std::string(who) + ": no group " + path.substr(0, j)};
                if(n->kids.size() >= kMaxMembers)
                    throw Error{std::string(who) + ": more than 256 members in a group"};
                next = new WNode();
                next->name = part;
                n->kids.push_back(next);
            }
            n = next;
            i = j;
        }
        return n;
    }
};

namespace {

int check_type(int cls, int size) {
    if(cls == 1) return size == 4 || size == 8;
    if(cls == 0 || cls == 2) return size == 1 || size == 2 || size == 4 || size == 8;
    return 0;
}

}  // namespace

extern "C" {

int sfx_h5w_create(sfx_h5w** out) {
    if(!out) return set_error(SFX_ERR_ARG, "sfx_h5w_create: NULL");
    *out = new sfx_h5w();
    return SFX_OK;
}

int sfx_h5w_destroy(sfx_h5w* w) {
    delete w;
    return SFX_OK;
}

int sfx_h5w_group(sfx_h5w* w, const char* path) {
    if(!w || !path) return set_error(SFX_ERR_ARG, "sfx_h5w_group: NULL");
    try {
        WNode* n = w->find(path, true, "sfx_h5w_group");
        if(!n->group) return set_error(SFX_ERR_ARG, "sfx_h5w_group: %s is a dataset", path);
        return SFX_OK;
    } catch(const Error& e) {
        return h5_fail("sfx_h5w_group", e);
    }
}

int sfx_h5w_dataset(sfx_h5w* w, const char* path, int type_class, int type_size, int rank,
                    const int64_t* dims, const void* data) {
    if(!w || !path || rank < 0 || rank > 8 || (rank && !dims))
        return set_error(SFX_ERR_ARG, "sfx_h5w_dataset: bad arguments");
    if(!check_type(type_class, type_size))
        return set_error(SFX_ERR_ARG, "sfx_h5w_dataset: type %d of %d bytes", type_class, type_size);
    try {
        std::string p(path);
        const size_t cut = p.find_last_of('/');
        WNode* parent = w->find(cut == std::string::npos ? "/" : p.substr(0, cut), false,
                                "sfx_h5w_dataset");
        const std::string name = cut == std::string::npos ? p : p.substr(cut + 1);
        if(name.empty() || !parent->group)
            return set_error(SFX_ERR_ARG, "sfx_h5w_dataset: bad path %s", path);
        for(WNode* k : parent->kids)
            if(k->name == name) return set_error(SFX_ERR_ARG, "sfx_h5w_dataset: %s exists", path);
        if(parent->kids.size() >= kMaxMembers)
            return set_error(SFX_ERR_ARG, "sfx_h5w_dataset: more than 256 members in a group");
        uint64_t count = 1;
        for(int d = 0; d < rank; d++) {
            if(dims[d] < 0) return set_error(SFX_ERR_ARG, "sfx_h5w_dataset: negative dimension");
            count *= (uint64_t)dims[d];
        }
        if(count && !data) return set_error(SFX_ERR_ARG, "sfx_h5w_dataset: NULL data");
        WNode* n = new WNode();
        n->name = name;
        n->group = false;
        n->cls = type_class;
        n->size = type_size;
        for(int d = 0; d < rank; d++) n->dims.push_back((uint64_t)dims[d]);
        n->data.assign((const uint8_t*)data, (const uint8_t*)data + count * (uint64_t)type_size);
        parent->kids.push_back(n);
        return SFX_OK;
    } catch(const Error& e) {
        return h5_fail("sfx_h5w_dataset", e);
    }
}

int sfx_h5w_attr(sfx_h5w* w, const char* path, const char* name, int type_class, int type_size,
                 int rank, const int64_t* dims, const void* data) {
    if(!w || !path || !name || !*name || rank < 0 || rank > 8 || (rank && !dims) || !data)
        return set_error(SFX_ERR_ARG, "sfx_h5w_attr: bad arguments");
    if(!check_type(type_class, type_size))
        return set_error(SFX_ERR_ARG, "sfx_h5w_attr: type %d of %d bytes", type_class, type_size);
    try {
        WNode* n = w->find(path, false, "sfx_h5w_attr");
        for(auto& a : n->attrs)
            if(a.name == name) return set_error(SFX_ERR_ARG, "sfx_h5w_attr: %s exists", name);
        WNode::WAttr a;
        a.name = name;
        a.cls = type_class;
        a.size = type_size;
        uint64_t count = 1;
        for(int d = 0; d < rank; d++) {
            a.dims.push_back((uint64_t)dims[d]);
            count *= (uint64_t)dims[d];
        }
        if(count * (uint64_t)type_size > 60000)
            return set_error(SFX_ERR_ARG, "sfx_h5w_attr: attribute too large");
        a.data.assign((const uint8_t*)data, (const uint8_t*)data + count * (uint64_t)type_size);
        n->attrs.push_back(a);
        return SFX_OK;
    } catch(const Error& e) {
        return h5_fail("sfx_h5w_attr", e);
    }
}

int sfx_h5w_save(sfx_h5w* w, const char* path) {
    if(!w || !path) return set_error(SFX_ERR_ARG, "sfx_h5w_save: NULL");
    uint64_t at = 96;  // superblock with the root symbol table entry
    place(&w->root, at);
    Out o;
    o.b.assign(at, 0);  // every structure's padding and unused slots stay zero
    static const uint8_t sig[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
    o.bytes(0, sig, 8);
    o.at(13, 8, 1);  // size of offsets
    o.at(14, 8, 1);  // size of lengths
    o.at(16, kLeafK, 2);
    o.at(18, kInternalK, 2);
    o.at(24, 0, 8);        // base address
    o.at(32, kUndef, 8);   // free-space info
    o.at(40, at, 8);       // end of file
    o.at(48, kUndef, 8);   // driver information
    o.at(56, 0, 8);        // root entry: name offset
    o.at(64, w->root.oh, 8);
    o.at(72, 1, 4);        // cached symbol table
    o.at(80, w->root.btree, 8);
    o.at(88, w->root.heap, 8);
    emit(&w->root, o);
    FILE* fp = std::fopen(path, "wb");
    if(!fp) return set_error(SFX_ERR_ARG, "sfx_h5w_save: cannot create %s", path);
    const size_t put = std::fwrite(o.b.data(), 1, o.b.size(), fp);
    const int closed = std::fclose(fp);
    if(put != o.b.size() || closed != 0)
        return set_error(SFX_ERR_ARG, "sfx_h5w_save: short write of %s", path);
    return SFX_OK;
}

}  // extern "C"
