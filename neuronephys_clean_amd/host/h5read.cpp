// h5read.cpp - a minimal reader of HDF5 datasets (no HDF5 library), and the
// reference's file loaders built on it: connectivity (/row_ptr, /col_idx,
// src/network_factory.cu:171-212), morphology (/x, /y, /z, /region_id,
// src_wholebrain/morphology.cpp:301-370) and the text CSR format of
// Csr_matrix::input_from_connection_file (src/paracrine.cuh:49-108).
//
// The reader follows the published HDF5 file format specification (version
// 3.0): superblocks 0-3, object headers 1 and 2 with continuation blocks,
// groups stored as symbol tables (version-1 B-tree + local heap) or as link
// messages in the header (compact storage), datasets of fixed-point or IEEE
// floating-point elements of either byte order in compact, contiguous or
// chunked storage (chunk indexes: version-1 B-tree, single chunk, implicit,
// fixed array), with the deflate, shuffle and fletcher32 filters.  Dense link
// storage (fractal heaps), extensible-array and version-2 B-tree chunk
// indexes, variable-length and compound types are reported as unsupported.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nnx_host.h"

namespace {

struct H5Error : std::runtime_error {
    explicit H5Error(const std::string& m) : std::runtime_error(m) {}
};

[[noreturn]] void fail(const char* fmt, ...) {
    char b[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(b, sizeof b, fmt, ap);
    va_end(ap);
    throw H5Error(b);
}

constexpr uint64_t kUndef = ~0ull;

struct Msg {
    int type;
    uint64_t pos, size;
};

struct Type {
    int cls = -1;  // 0 fixed-point, 1 floating-point
    int size = 0;
    bool be = false, sgn = false;
};

struct Filter {
    int id;
    int flags;
    std::vector<uint32_t> cd;
};

class File {
    const uint8_t* d_ = nullptr;
    size_t n_ = 0;
    int fd_ = -1;

   public:
    int so = 8, sl = 8;  // sizes of offsets and lengths
    uint64_t base = 0, root = kUndef;
    uint64_t root_btree = kUndef, root_heap = kUndef;  // superblock 0/1 root symbol table (cached)

    explicit File(const char* path) {
        fd_ = ::open(path, O_RDONLY);
        if (fd_ < 0) fail("cannot open %s", path);
        struct stat st;
        if (fstat(fd_, &st) != 0 || st.st_size <= 0) fail("cannot stat %s", path);
        n_ = (size_t)st.st_size;
        void* p = mmap(nullptr, n_, PROT_READ, MAP_PRIVATE, fd_, 0);
        if (p == MAP_FAILED) fail("cannot map %s", path);
        d_ = (const uint8_t*)p;
        superblock();
    }
    ~File() {
        if (d_) munmap((void*)d_, n_);
        if (fd_ >= 0) ::close(fd_);
    }
    File(const File&) = delete;
    File& operator=(const File&) = delete;

    void need(uint64_t p, uint64_t k) const {
        if (p > n_ || k > n_ - p) fail("read past the end of the file (offset %llu)", (unsigned long long)p);
    }
    const uint8_t* at(uint64_t p, uint64_t k) const {
        need(p, k);
        return d_ + p;
    }
    uint64_t uint(uint64_t p, int k) const {  // little-endian unsigned of k bytes
        const uint8_t* b = at(p, k);
        uint64_t v = 0;
        for (int i = k - 1; i >= 0; --i) v = (v << 8) | b[i];
        return v;
    }
    uint64_t addr(uint64_t p) const {
        const uint64_t v = uint(p, so);
        if (so < 8 && v == (1ull << (8 * so)) - 1) return kUndef;
        return v == kUndef ? kUndef : v + base;
    }
    uint64_t len(uint64_t p) const { return uint(p, sl); }
    bool sig(uint64_t p, const char* s) const { return p + 4 <= n_ && std::memcmp(d_ + p, s, 4) == 0; }

    void superblock() {
        static const uint8_t magic[8] = {0x89, 'H', 'D', 'F', '\r', '\n', 0x1a, '\n'};
        uint64_t sb = kUndef;
        for (uint64_t p = 0; p + 8 <= n_; p = p ? 2 * p : 512)
            if (std::memcmp(d_ + p, magic, 8) == 0) {
                sb = p;
                break;
            }
        if (sb == kUndef) fail("not an HDF5 file (no superblock signature)");
        const int ver = (int)uint(sb + 8, 1);
        if (ver <= 1) {
            so = (int)uint(sb + 13, 1);
            sl = (int)uint(sb + 14, 1);
            uint64_t p = sb + 24 + (ver == 1 ? 4 : 0);
            base = uint(p, so);
            p += 4 * (uint64_t)so;  // base, free space, end of file, driver block
            // root group symbol table entry: name offset, header address,
            // cache type, reserved, scratch (B-tree and heap addresses)
            root = addr(p + so);
            const uint32_t cache = (uint32_t)uint(p + 2 * so, 4);
            if (cache == 1) {
                root_btree = addr(p + 2 * so + 8);
                root_heap = addr(p + 3 * so + 8);
            }
        } else if (ver <= 3) {
            so = (int)uint(sb + 9, 1);
            sl = (int)uint(sb + 10, 1);
            base = uint(sb + 12, so);
            root = addr(sb + 12 + 3 * (uint64_t)so);
        } else {
            fail("unsupported superblock version %d", ver);
        }
        if (so != 2 && so != 4 && so != 8) fail("bad size of offsets %d", so);
        if (sl != 2 && sl != 4 && sl != 8) fail("bad size of lengths %d", sl);
    }

    // All messages of an object header, following continuation blocks.
    std::vector<Msg> messages(uint64_t oh) const {
        std::vector<Msg> out;
        std::vector<std::pair<uint64_t, uint64_t>> blocks;  // (start, end) of message areas
        bool v2 = sig(oh, "OHDR");
        bool order = false;
        if (v2) {
            const int flags = (int)uint(oh + 5, 1);
            order = flags & 0x04;
            uint64_t p = oh + 6 + ((flags & 0x20) ? 16 : 0) + ((flags & 0x10) ? 4 : 0);
            const int csz = 1 << (flags & 3);
            const uint64_t size0 = uint(p, csz);
            p += csz;
            blocks.push_back({p, p + size0});
        } else {
            if (uint(oh, 1) != 1) fail("unknown object header version at %llu", (unsigned long long)oh);
            const uint64_t hsize = uint(oh + 8, 4);
            blocks.push_back({oh + 16, oh + 16 + hsize});
        }
        for (size_t b = 0; b < blocks.size(); ++b) {
            uint64_t p = blocks[b].first, end = blocks[b].second;
            if (v2 && b > 0) {  // continuation block: signature, messages, checksum
                if (!sig(p, "OCHK")) fail("bad continuation block");
                p += 4;
                end -= 4;
            }
            const uint64_t hdr = v2 ? 4 + (order ? 2 : 0) : 8;
            while (p + hdr <= end) {
                Msg m;
                if (v2) {
                    m.type = (int)uint(p, 1);
                    m.size = uint(p + 1, 2);
                } else {
                    m.type = (int)uint(p, 2);
                    m.size = uint(p + 2, 2);
                }
                m.pos = p + hdr;
                need(m.pos, m.size);
                if (m.type == 0x10) {
                    if (blocks.size() > 4096) fail("too many object header continuation blocks");
                    blocks.push_back({addr(m.pos), addr(m.pos) + len(m.pos + so)});
                }
                else out.push_back(m);
                p = m.pos + m.size;
            }
        }
        return out;
    }

    // Object header address of `name` inside the group whose header is `oh`.
    uint64_t child(uint64_t oh, const std::string& name, uint64_t btree = kUndef, uint64_t heap = kUndef) const {
        const std::vector<Msg> ms = messages(oh);
        for (const Msg& m : ms)
            if (m.type == 0x11) {
                btree = addr(m.pos);
                heap = addr(m.pos + so);
            }
        if (btree != kUndef) return symtab_find(btree, heap, name);
        for (const Msg& m : ms)
            if (m.type == 0x06) {
                std::string nm;
                uint64_t target;
                if (link(m, nm, target) && nm == name) return target;
            }
        for (const Msg& m : ms)
            if (m.type == 0x02) {
                const int flags = (int)uint(m.pos + 1, 1);
                const uint64_t fh = addr(m.pos + 2 + ((flags & 1) ? 8 : 0));
                if (fh != kUndef) fail("group uses dense link storage (not supported): looking for %s", name.c_str());
            }
        fail("no object named %s", name.c_str());
    }

    // A link message: its name and, for a hard link, the target header.
    bool link(const Msg& m, std::string& name, uint64_t& target) const {
        uint64_t p = m.pos;
        if (uint(p, 1) != 1) fail("unknown link message version");
        const int flags = (int)uint(p + 1, 1);
        p += 2;
        int ltype = 0;
        if (flags & 0x08) ltype = (int)uint(p++, 1);
        if (flags & 0x04) p += 8;
        if (flags & 0x10) p += 1;
        const int lk = 1 << (flags & 3);
        const uint64_t nl = uint(p, lk);
        p += lk;
        name.assign((const char*)at(p, nl), nl);
        p += nl;
        if (ltype != 0) return false;  // soft / external links are not followed
        target = addr(p);
        return true;
    }

    uint64_t symtab_find(uint64_t node, uint64_t heap, const std::string& name) const {
        if (!sig(heap, "HEAP")) fail("bad local heap");
        const uint64_t data = addr(heap + 8 + 2 * (uint64_t)sl);
        uint64_t found = kUndef;
        symtab_walk(node, data, name, found, 0);
        if (found == kUndef) fail("no object named %s", name.c_str());
        return found;
    }
    void symtab_walk(uint64_t node, uint64_t heap_data, const std::string& name, uint64_t& found, int depth) const {
        if (depth > 64) fail("group B-tree too deep");
        if (!sig(node, "TREE") || uint(node + 4, 1) != 0) fail("bad group B-tree node");
        const int level = (int)uint(node + 5, 1);
        const int used = (int)uint(node + 6, 2);
        const uint64_t p0 = node + 8 + 2 * (uint64_t)so;  // past the sibling addresses
        for (int i = 0; i < used && found == kUndef; ++i) {
            const uint64_t ch = addr(p0 + (uint64_t)sl + (uint64_t)i * (sl + so));
            if (level > 0) {
                symtab_walk(ch, heap_data, name, found, depth + 1);
                continue;
            }
            if (!sig(ch, "SNOD")) fail("bad symbol table node");
            const int nsym = (int)uint(ch + 6, 2);
            for (int s = 0; s < nsym; ++s) {
                const uint64_t e = ch + 8 + (uint64_t)s * (2 * so + 24);
                const uint64_t off = uint(e, so);
                const char* nm = (const char*)at(heap_data + off, 1);
                const size_t maxl = n_ - (heap_data + off);
                if (strnlen(nm, maxl) == name.size() && std::memcmp(nm, name.data(), name.size()) == 0) {
                    found = addr(e + so);
                    return;
                }
            }
        }
    }

    uint64_t resolve(const std::string& path) const {
        if (root == kUndef) fail("no root group");
        uint64_t oh = root;
        size_t i = 0;
        bool first = true;
        while (i < path.size()) {
            while (i < path.size() && path[i] == '/') ++i;
            if (i >= path.size()) break;
            size_t j = path.find('/', i);
            if (j == std::string::npos) j = path.size();
            const std::string comp = path.substr(i, j - i);
            oh = first && root_btree != kUndef ? child(oh, comp, root_btree, root_heap) : child(oh, comp);
            first = false;
            i = j;
        }
        return oh;
    }

    // ------------------------------------------------------------ datasets
    struct Dataset {
        std::vector<uint64_t> dims;
        Type t;
        int layout = -1;  // 0 compact, 1 contiguous, 2 chunked
        int lver = 0;
        uint64_t data = kUndef, dsize = 0;  // compact / contiguous; chunked: index address
        std::vector<uint64_t> cdims;        // chunk dims (rank)
        int index = 0;                      // 0: v1 B-tree (layout 3); 1-5: layout 4 index types
        int cflags = 0;
        uint64_t single_fsize = 0;
        uint32_t single_mask = 0;
        std::vector<Filter> filters;
        uint64_t count() const {
            uint64_t c = 1;
            for (uint64_t x : dims) c *= x;
            return c;
        }
    };

    Dataset dataset(uint64_t oh) const {
        Dataset ds;
        bool space = false;
        for (const Msg& m : messages(oh)) {
            const uint64_t p = m.pos;
            if (m.type == 0x01) {
                space = true;
                const int ver = (int)uint(p, 1);
                const int rank = (int)uint(p + 1, 1);
                uint64_t q;
                if (ver == 1) {
                    q = p + 8;
                } else if (ver == 2) {
                    if (uint(p + 3, 1) == 2) fail("null dataspace");
                    q = p + 4;
                } else {
                    fail("unknown dataspace version %d", ver);
                }
                for (int r = 0; r < rank; ++r) ds.dims.push_back(len(q + (uint64_t)r * sl));
            } else if (m.type == 0x03) {
                const int cls = (int)(uint(p, 1) & 0x0f);
                const uint32_t bits = (uint32_t)uint(p + 1, 3);
                ds.t.size = (int)uint(p + 4, 4);
                if (cls == 0) {
                    ds.t.cls = 0;
                    ds.t.be = bits & 1;
                    ds.t.sgn = bits & 8;
                    if (ds.t.size != 1 && ds.t.size != 2 && ds.t.size != 4 && ds.t.size != 8)
                        fail("unsupported integer size %d", ds.t.size);
                } else if (cls == 1) {
                    ds.t.cls = 1;
                    if (bits & 0x40) fail("VAX floating-point order is not supported");
                    ds.t.be = bits & 1;
                    if (ds.t.size != 4 && ds.t.size != 8) fail("unsupported float size %d", ds.t.size);
                } else {
                    fail("unsupported datatype class %d (only integers and IEEE floats)", cls);
                }
            } else if (m.type == 0x08) {
                layout_msg(m, ds);
            } else if (m.type == 0x0b) {
                filter_msg(m, ds);
            }
        }
        if (!space || ds.t.cls < 0 || ds.layout < 0) fail("object is not a dataset of numbers");
        return ds;
    }

    void layout_msg(const Msg& m, Dataset& ds) const {
        const uint64_t p = m.pos;
        ds.lver = (int)uint(p, 1);
        if (ds.lver < 3 || ds.lver > 4) fail("unsupported layout message version %d", ds.lver);
        ds.layout = (int)uint(p + 1, 1);
        if (ds.layout == 0) {
            ds.dsize = uint(p + 2, 2);
            ds.data = p + 4;
        } else if (ds.layout == 1) {
            ds.data = addr(p + 2);
            ds.dsize = len(p + 2 + so);
        } else if (ds.layout == 2 && ds.lver == 3) {
            const int dimn = (int)uint(p + 2, 1);
            ds.data = addr(p + 3);
            for (int r = 0; r + 1 < dimn; ++r) ds.cdims.push_back(uint(p + 3 + so + 4 * (uint64_t)r, 4));
            ds.index = 0;
        } else if (ds.layout == 2) {
            ds.cflags = (int)uint(p + 2, 1);
            const int dimn = (int)uint(p + 3, 1);
            const int enc = (int)uint(p + 4, 1);
            uint64_t q = p + 5;
            for (int r = 0; r < dimn; ++r, q += enc)
                if (r + 1 < dimn) ds.cdims.push_back(uint(q, enc));
            ds.index = (int)uint(q, 1);
            q += 1;
            if (ds.index == 1) {
                if (ds.cflags & 2) {
                    ds.single_fsize = len(q);
                    ds.single_mask = (uint32_t)uint(q + sl, 4);
                    q += sl + 4;
                }
            } else if (ds.index == 3) {
                q += 1;
            } else if (ds.index != 2) {
                fail("unsupported chunk index type %d (extensible array / version-2 B-tree)", ds.index);
            }
            ds.data = addr(q);
        } else {
            fail("unsupported storage layout class %d", ds.layout);
        }
    }

    void filter_msg(const Msg& m, Dataset& ds) const {
        uint64_t p = m.pos;
        const int ver = (int)uint(p, 1);
        const int nf = (int)uint(p + 1, 1);
        p += ver == 1 ? 8 : 2;
        for (int f = 0; f < nf; ++f) {
            Filter fl;
            fl.id = (int)uint(p, 2);
            uint64_t namelen = 0;
            if (ver == 1 || fl.id >= 256) {
                namelen = uint(p + 2, 2);
                p += 4;
            } else {
                p += 2;
            }
            fl.flags = (int)uint(p, 2);
            const int nv = (int)uint(p + 2, 2);
            p += 4;
            if (ver == 1) namelen = (namelen + 7) / 8 * 8;
            p += namelen;
            for (int v = 0; v < nv; ++v) fl.cd.push_back((uint32_t)uint(p + 4 * (uint64_t)v, 4));
            p += 4 * (uint64_t)nv;
            if (ver == 1 && (nv & 1)) p += 4;
            if (fl.id != 1 && fl.id != 2 && fl.id != 3) fail("unsupported filter %d", fl.id);
            ds.filters.push_back(fl);
        }
    }

    // One stored chunk, filters undone, into `out` (chunk bytes).
    void chunk_bytes(const Dataset& ds, uint64_t a, uint64_t stored, uint32_t mask, std::vector<uint8_t>& out) const {
        const uint64_t es = ds.t.size;
        uint64_t cb = es;
        for (uint64_t c : ds.cdims) cb *= c;
        std::vector<uint8_t> buf(at(a, stored), at(a, stored) + stored);
        for (int f = (int)ds.filters.size() - 1; f >= 0; --f) {
            if (mask & (1u << f)) continue;
            const Filter& fl = ds.filters[f];
            if (fl.id == 1) {
                std::vector<uint8_t> o(cb);
                uLongf ol = (uLongf)cb;
                if (uncompress(o.data(), &ol, buf.data(), (uLong)buf.size()) != Z_OK || ol != cb)
                    fail("deflate: corrupt chunk");
                buf.swap(o);
            } else if (fl.id == 2) {
                const uint64_t w = fl.cd.empty() ? es : fl.cd[0];
                const uint64_t ne = buf.size() / w;
                std::vector<uint8_t> o(buf.size());
                for (uint64_t b = 0; b < w; ++b)
                    for (uint64_t e = 0; e < ne; ++e) o[e * w + b] = buf[b * ne + e];
                std::copy(buf.begin() + ne * w, buf.end(), o.begin() + ne * w);
                buf.swap(o);
            } else if (fl.id == 3) {
                if (buf.size() < 4) fail("fletcher32: short chunk");
                buf.resize(buf.size() - 4);
            }
        }
        if (buf.size() < cb) fail("chunk smaller than its dimensions");
        buf.resize(cb);
        out.swap(buf);
    }

    // Copies one chunk (at chunk offsets `off`) into the row-major dataset.
    static void place(const Dataset& ds, const std::vector<uint64_t>& off, const std::vector<uint8_t>& chunk,
                      uint8_t* dst) {
        const int rank = (int)ds.dims.size();
        const uint64_t es = ds.t.size;
        if (rank == 0) {
            std::memcpy(dst, chunk.data(), es);
            return;
        }
        std::vector<uint64_t> idx(rank, 0);
        const uint64_t rowlen = std::min(ds.cdims[rank - 1], ds.dims[rank - 1] - std::min(off[rank - 1], ds.dims[rank - 1]));
        if (off[rank - 1] >= ds.dims[rank - 1]) return;
        for (;;) {
            bool inside = true;
            uint64_t d = 0, s = 0;
            for (int r = 0; r < rank; ++r) {
                const uint64_t g = off[r] + idx[r];
                if (g >= ds.dims[r]) inside = false;
                d = d * ds.dims[r] + g;
                s = s * ds.cdims[r] + idx[r];
            }
            if (inside) std::memcpy(dst + d * es, chunk.data() + s * es, rowlen * es);
            int r = rank - 2;
            for (; r >= 0; --r) {
                if (++idx[r] < ds.cdims[r]) break;
                idx[r] = 0;
            }
            if (r < 0) break;
        }
    }

    void btree_chunks(const Dataset& ds, uint64_t node, uint8_t* dst, int depth) const {
        if (depth > 64) fail("chunk B-tree too deep");
        if (!sig(node, "TREE") || uint(node + 4, 1) != 1) fail("bad chunk B-tree node");
        const int level = (int)uint(node + 5, 1);
        const int used = (int)uint(node + 6, 2);
        const int rank = (int)ds.dims.size();
        const uint64_t ksz = 8 + 8 * (uint64_t)(rank + 1);
        const uint64_t p0 = node + 8 + 2 * (uint64_t)so;
        std::vector<uint8_t> chunk;
        for (int i = 0; i < used; ++i) {
            const uint64_t key = p0 + (uint64_t)i * (ksz + so);
            const uint64_t ch = addr(key + ksz);
            if (level > 0) {
                btree_chunks(ds, ch, dst, depth + 1);
                continue;
            }
            const uint64_t stored = uint(key, 4);
            const uint32_t mask = (uint32_t)uint(key + 4, 4);
            std::vector<uint64_t> off(rank);
            for (int r = 0; r < rank; ++r) off[r] = uint(key + 8 + 8 * (uint64_t)r, 8);
            chunk_bytes(ds, ch, stored, mask, chunk);
            place(ds, off, chunk, dst);
        }
    }

    // offsets of chunk number i in row-major order over the chunk grid
    static std::vector<uint64_t> chunk_offsets(const Dataset& ds, uint64_t i) {
        const int rank = (int)ds.dims.size();
        std::vector<uint64_t> off(rank);
        for (int r = rank - 1; r >= 0; --r) {
            const uint64_t nc = (ds.dims[r] + ds.cdims[r] - 1) / ds.cdims[r];
            off[r] = (i % nc) * ds.cdims[r];
            i /= nc;
        }
        return off;
    }

    void read(const Dataset& ds, uint8_t* dst) const {
        const uint64_t es = ds.t.size, total = ds.count() * es;
        std::memset(dst, 0, total);  // unallocated storage reads as the default fill value 0
        if (ds.layout == 0 || ds.layout == 1) {
            if (!ds.filters.empty()) fail("filters on unchunked storage");
            if (ds.data == kUndef) return;
            if (ds.dsize < total) fail("dataset storage smaller than its dataspace");
            std::memcpy(dst, at(ds.data, total), total);
            return;
        }
        if (ds.cdims.size() != ds.dims.size()) fail("chunk rank does not match the dataspace");
        for (uint64_t c : ds.cdims)
            if (c == 0) fail("zero chunk dimension");
        if (ds.data == kUndef) return;
        uint64_t nchunks = 1, cb = es;
        for (size_t r = 0; r < ds.dims.size(); ++r) {
            nchunks *= (ds.dims[r] + ds.cdims[r] - 1) / ds.cdims[r];
            cb *= ds.cdims[r];
        }
        std::vector<uint8_t> chunk;
        if (ds.lver == 3 || ds.index == 0) {
            btree_chunks(ds, ds.data, dst, 0);
        } else if (ds.index == 1) {
            chunk_bytes(ds, ds.data, (ds.cflags & 2) ? ds.single_fsize : cb, ds.single_mask, chunk);
            place(ds, std::vector<uint64_t>(ds.dims.size(), 0), chunk, dst);
        } else if (ds.index == 2) {
            for (uint64_t i = 0; i < nchunks; ++i) {
                chunk_bytes(ds, ds.data + i * cb, cb, 0, chunk);
                place(ds, chunk_offsets(ds, i), chunk, dst);
            }
        } else if (ds.index == 3) {
            const uint64_t h = ds.data;
            if (!sig(h, "FAHD")) fail("bad fixed array header");
            const int client = (int)uint(h + 5, 1);
            const int esz = (int)uint(h + 6, 1);
            const int pbits = (int)uint(h + 7, 1);
            const uint64_t nel = len(h + 8);
            const uint64_t db = addr(h + 8 + sl);
            if (nel > (1ull << pbits)) fail("paged fixed array chunk index is not supported");
            if (db == kUndef) return;
            if (!sig(db, "FADB")) fail("bad fixed array data block");
            const uint64_t e0 = db + 6 + so;
            for (uint64_t i = 0; i < nel && i < nchunks; ++i) {
                const uint64_t e = e0 + i * esz;
                const uint64_t a = addr(e);
                if (a == kUndef) continue;
                uint64_t stored = cb;
                uint32_t mask = 0;
                if (client == 1) {
                    const int sz = esz - so - 4;
                    stored = uint(e + so, sz);
                    mask = (uint32_t)uint(e + so + sz, 4);
                }
                chunk_bytes(ds, a, stored, mask, chunk);
                place(ds, chunk_offsets(ds, i), chunk, dst);
            }
        }
    }
};

template <class O> void convert(const Type& t, const uint8_t* src, uint64_t n, O* out) {
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t* b = src + i * t.size;
        uint64_t u = 0;
        for (int k = 0; k < t.size; ++k) u |= (uint64_t)b[t.be ? t.size - 1 - k : k] << (8 * k);
        if (t.cls == 1) {
            if (t.size == 4) {
                float f;
                uint32_t w = (uint32_t)u;
                std::memcpy(&f, &w, 4);
                out[i] = (O)f;
            } else {
                double f;
                std::memcpy(&f, &u, 8);
                out[i] = (O)f;
            }
        } else if (t.sgn) {
            const int sh = 64 - 8 * t.size;
            out[i] = (O)((int64_t)(u << sh) >> sh);
        } else {
            out[i] = (O)u;
        }
    }
}

void put_err(char* err, size_t errlen, const char* m) {
    if (err && errlen) {
        std::strncpy(err, m, errlen - 1);
        err[errlen - 1] = 0;
    }
}

}  // namespace

extern "C" {

int64_t nnx_h5_count(const char* path, const char* name, char* err, size_t errlen) {
    try {
        File f(path);
        return (int64_t)f.dataset(f.resolve(name)).count();
    } catch (const std::exception& e) {
        put_err(err, errlen, e.what());
        return -1;
    }
}

int64_t nnx_h5_read(const char* path, const char* name, int out_type, void* out, int64_t cap, char* err,
                    size_t errlen) {
    try {
        File f(path);
        const File::Dataset ds = f.dataset(f.resolve(name));
        const uint64_t n = ds.count();
        if ((int64_t)n > cap) fail("dataset %s has %llu elements, room for %lld", name, (unsigned long long)n,
                                   (long long)cap);
        std::vector<uint8_t> raw(n * ds.t.size + 1);
        f.read(ds, raw.data());
        switch (out_type) {
            case NNX_H5_I32: convert(ds.t, raw.data(), n, (int32_t*)out); break;
            case NNX_H5_I64: convert(ds.t, raw.data(), n, (int64_t*)out); break;
            case NNX_H5_F32: convert(ds.t, raw.data(), n, (float*)out); break;
            case NNX_H5_F64: convert(ds.t, raw.data(), n, (double*)out); break;
            default: fail("bad output type %d", out_type);
        }
        return (int64_t)n;
    } catch (const std::exception& e) {
        put_err(err, errlen, e.what());
        return -1;
    }
}

int64_t nnx_csr_subsample(int64_t numpnts, const int32_t* row_ptr, const int32_t* col_idx, int64_t N,
                          int32_t* offsets, int32_t* targets, int64_t* dropped) {
    if (N < 1 || numpnts < N) return -1;
    const int64_t sparsity = numpnts / N;
    int64_t k = 0, drop = 0;
    offsets[0] = 0;
    for (int64_t from = 0; from < N; ++from) {
        const int64_t r = from * sparsity;
        for (int64_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
            const int32_t c = col_idx[e];
            if (c < 0 || c % sparsity != 0) continue;
            if (c / sparsity >= N) {
                ++drop;
                continue;
            }
            targets[k++] = (int32_t)(c / sparsity);
        }
        offsets[from + 1] = (int32_t)k;
    }
    if (dropped) *dropped = drop;
    return k;
}

int nnx_morphology_from_axes(int64_t numpnts, const float* x, const float* y, const float* z, int64_t N,
                             float* pos) {
    if (N < 1 || numpnts < N) return -1;
    const float* ax[3] = {x, y, z};
    float lo[3], hi[3], sc[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = hi[d] = ax[d][0];
        for (int64_t i = 1; i < numpnts; ++i) {
            lo[d] = std::min(lo[d], ax[d][i]);
            hi[d] = std::max(hi[d], ax[d][i]);
        }
        sc[d] = 2.0f / (hi[d] - lo[d]);
    }
    const int64_t stride = numpnts / N;
    for (int64_t i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) {
            float p = ax[d][stride * i];
            p = p + (-lo[d]);
            p = p * sc[d];
            p = p + (-1.0f);
            pos[3 * i + d] = p;
        }
    return 0;
}

int64_t nnx_read_connection_file(const char* path, int64_t* rows, int64_t* nnz, int32_t* row_ptr, int32_t* col,
                                 float* val, char* err, size_t errlen) {
    FILE* f = std::fopen(path, "r");
    if (!f) {
        put_err(err, errlen, "cannot open the connection file");
        return -1;
    }
    std::string text;
    char buf[1 << 16];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, k);
    std::fclose(f);
    std::vector<std::string> lines;
    size_t s = 0;
    while (s <= text.size() && lines.size() < 4) {
        size_t e = text.find('\n', s);
        if (e == std::string::npos) e = text.size();
        lines.push_back(text.substr(s, e - s));
        s = e + 1;
    }
    long long n = -1, m = -1;
    if (lines.empty() || std::sscanf(lines[0].c_str(), "%lld %lld", &n, &m) != 2 || n < 0 || m < 0) {
        put_err(err, errlen, "bad header line (expected: neuronCount nnz)");
        return -1;
    }
    *rows = n;
    *nnz = m;
    if (!row_ptr) return 0;  // sizes only
    if (lines.size() < 4) {
        put_err(err, errlen, "expected four lines: header, rowPtr, colIdx, val");
        return -1;
    }
    auto ints = [&](const std::string& l, int64_t cnt, int32_t* o) {
        const char* p = l.c_str();
        for (int64_t i = 0; i < cnt; ++i) {
            char* q;
            const long v = std::strtol(p, &q, 10);
            if (q == p) return false;
            o[i] = (int32_t)v;
            p = q;
        }
        return true;
    };
    if (!ints(lines[1], n + 1, row_ptr) || !ints(lines[2], m, col)) {
        put_err(err, errlen, "short rowPtr or colIdx line");
        return -1;
    }
    const char* p = lines[3].c_str();
    for (int64_t i = 0; i < m; ++i) {
        char* q;
        const float v = std::strtof(p, &q);
        if (q == p) {
            put_err(err, errlen, "short val line");
            return -1;
        }
        val[i] = v;
        p = q;
    }
    return 0;
}

}  // extern "C"
