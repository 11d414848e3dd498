// Checkpoint and restore of a basin's physical state (the reference's
// save_sparse_checkpoint / load_sparse_checkpoint, checkpoint.jl:7-66), in a versioned raw
// format of this library instead of Julia's version-locked `Serialization`:
//
//   offset 0   "MAELCKPT"                      8 bytes
//          8   u32 format version (1), u32 kind (0 dense, 1 sparse)
//         16   u64 length of the metadata, then that many bytes of UTF-8 JSON written by
//              the host (the basin, grid, stirring model, clocks: what the reference keeps
//              on its host)
//          .   u32 number of arrays, u32 0, then per array a 64-byte record:
//              char name[32] (NUL-padded), u32 element size, u32 0, u64 count,
//              u64 byte offset of its data from the file's start, u64 0
//          .   the arrays' data, each at an offset that is a multiple of 256
//
// All integers little-endian. The arrays are the state a step reads: the level set, types,
// face velocities and flags, pressure, the live particles (n of them) and, for the sparse
// basin, the topology (slot table, the live slots' blocks, the free list) with the fine
// pools cut to the live slots, so a checkpoint restores into a larger pool. φ0 and u0 are
// not stored: a restore sets them from φ and u, as the reference's loader does.
#include "capi_common.hpp"
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

constexpr char MAGIC[8] = {'M', 'A', 'E', 'L', 'C', 'K', 'P', 'T'};
constexpr uint32_t VERSION = 1;

struct Entry {
    std::string name;
    void* dev;                    // device pointer (or nullptr: `host` holds the data)
    int64_t count;                // elements
    int32_t es;                   // element size
    const void* host = nullptr;
};
struct Record {
    char name[32];
    uint32_t es, pad;
    uint64_t count, offset, reserved;
};
static_assert(sizeof(Record) == 64, "record layout");

uint64_t align256(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

struct File {
    FILE* f = nullptr;
    File(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode)) {
        if (!f) throw mael::Error(mael::E_ARG, "cannot open " + path);
    }
    ~File() { if (f) std::fclose(f); }
    void write(const void* p, size_t n) {
        if (n && std::fwrite(p, 1, n, f) != n) throw mael::Error(mael::E_ARG, "checkpoint write failed");
    }
    void read(void* p, size_t n) {
        if (n && std::fread(p, 1, n, f) != n) throw mael::Error(mael::E_ARG, "checkpoint truncated");
    }
    void seek(uint64_t off) {
        if (std::fseek(f, (long)off, SEEK_SET) != 0) throw mael::Error(mael::E_ARG, "checkpoint truncated");
    }
};

// Write `entries` (device arrays, copied through host staging on `s`) and the metadata to
// path + ".tmp", then rename it over `path`.
void write_checkpoint(const std::string& path, uint32_t kind, const std::string& meta, const std::vector<Entry>& entries,
                      hipStream_t s) {
    std::string tmp = path + ".tmp";
    {
        File out(tmp, "wb");
        uint64_t mlen = meta.size();
        uint32_t hdr[2] = {VERSION, kind};
        out.write(MAGIC, 8);
        out.write(hdr, 8);
        out.write(&mlen, 8);
        out.write(meta.data(), mlen);
        uint32_t na[2] = {(uint32_t)entries.size(), 0};
        out.write(na, 8);
        uint64_t off = align256(16 + 8 + mlen + 8 + 64 * (uint64_t)entries.size());
        std::vector<Record> recs(entries.size());
        for (size_t k = 0; k < entries.size(); ++k) {
            Record& r = recs[k];
            std::memset(&r, 0, sizeof r);
            std::strncpy(r.name, entries[k].name.c_str(), sizeof r.name - 1);
            r.es = (uint32_t)entries[k].es;
            r.count = (uint64_t)entries[k].count;
            r.offset = off;
            off = align256(off + r.count * r.es);
        }
        out.write(recs.data(), recs.size() * sizeof(Record));
        std::vector<char> stage;
        for (size_t k = 0; k < entries.size(); ++k) {
            out.seek(recs[k].offset);
            size_t n = recs[k].count * recs[k].es;
            if (entries[k].host) {
                out.write(entries[k].host, n);
                continue;
            }
            stage.resize(n);
            if (n) {
                MAEL_HIP(hipMemcpyAsync(stage.data(), entries[k].dev, n, hipMemcpyDeviceToHost, s));
                MAEL_HIP(hipStreamSynchronize(s));
            }
            out.write(stage.data(), n);
        }
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) throw mael::Error(mael::E_ARG, "cannot move the checkpoint into " + path);
}

struct Reader {
    uint32_t kind = 0;
    std::string meta;
    std::vector<Record> recs;
    std::unique_ptr<File> in;
    explicit Reader(const std::string& path) : in(new File(path, "rb")) {
        char magic[8];
        uint32_t hdr[2];
        uint64_t mlen;
        in->read(magic, 8);
        if (std::memcmp(magic, MAGIC, 8) != 0) throw mael::Error(mael::E_ARG, path + " is not a checkpoint");
        in->read(hdr, 8);
        if (hdr[0] != VERSION) throw mael::Error(mael::E_ARG, "Checkpoint format differs");
        kind = hdr[1];
        in->read(&mlen, 8);
        if (mlen > (1ull << 30)) throw mael::Error(mael::E_ARG, "checkpoint metadata too large");
        meta.resize(mlen);
        in->read(&meta[0], mlen);
        uint32_t na[2];
        in->read(na, 8);
        recs.resize(na[0]);
        in->read(recs.data(), recs.size() * sizeof(Record));
    }
    const Record& find(const std::string& name) const {
        for (auto& r : recs)
            if (name == std::string(r.name, strnlen(r.name, sizeof r.name))) return r;
        throw mael::Error(mael::E_ARG, "checkpoint has no array " + name);
    }
    template <class T> std::vector<T> host(const std::string& name) {
        const Record& r = find(name);
        if (r.es != sizeof(T)) throw mael::Error(mael::E_ARG, "checkpoint array " + name + " has the wrong element size");
        std::vector<T> v(r.count);
        in->seek(r.offset);
        in->read(v.data(), r.count * sizeof(T));
        return v;
    }
    // Copy array `name` into device memory of `count` elements of size `es` (it must fit).
    void upload(const std::string& name, void* dev, int64_t cap, int32_t es, hipStream_t s, int64_t want = -1) {
        const Record& r = find(name);
        if ((int32_t)r.es != es || (int64_t)r.count > cap || (want >= 0 && (int64_t)r.count != want))
            throw mael::Error(mael::E_ARG, "checkpoint array " + name + " does not fit this basin");
        std::vector<char> stage(r.count * r.es);
        in->seek(r.offset);
        in->read(stage.data(), stage.size());
        if (!stage.empty()) {
            MAEL_HIP(hipMemcpyAsync(dev, stage.data(), stage.size(), hipMemcpyHostToDevice, s));
            MAEL_HIP(hipStreamSynchronize(s));
        }
    }
};

const char* PNAMES[mael::PFIELDS] = {"px", "py", "pz", "vx", "vy", "vz", "c0", "c1", "c2", "c3", "c4", "c5", "c6", "c7",
                                     "c8", "lag"};

}  // namespace

extern "C" {

int mael_checkpoint_meta(const char* path, int32_t* kind, char* buf, int64_t cap, int64_t* len) {
    if (!path) return MAEL_E_ARG;
    try {
        Reader r(path);
        if (kind) *kind = (int32_t)r.kind;
        if (len) *len = (int64_t)r.meta.size();
        if (buf && cap > 0) {
            int64_t n = std::min<int64_t>(cap - 1, (int64_t)r.meta.size());
            std::memcpy(buf, r.meta.data(), (size_t)n);
            buf[n] = 0;
        }
        return MAEL_OK;
    } catch (const mael::Error& e) {
        return e.code;
    }
}

int mael_dense_checkpoint(mael_dense* d, const char* path, const char* meta) {
    if (!d || !path) return MAEL_E_ARG;
    return guard(d->ctx, [&] {
        mael::DenseBasin& b = *d->b;
        b.st.sync();
        const int64_t C = (int64_t)b.cells();
        std::vector<Entry> e = {{"phi", b.phi.p, C, 4}, {"types", b.types.p, C, 1}, {"P", b.P.p, C, 4}};
        const char* un[3] = {"u", "v", "w"};
        const char* fn[3] = {"fu", "fv", "fw"};
        for (int a = 0; a < 3; ++a) {
            e.push_back({un[a], b.u[a].p, (int64_t)b.faces(a), 4});
            e.push_back({fn[a], b.flag[a].p, (int64_t)b.faces(a), 1});
        }
        for (int f = 0; f < mael::PFIELDS; ++f) e.push_back({PNAMES[f], b.pbuf[b.cur][f].p, b.n, 4});
        write_checkpoint(path, 0, meta ? meta : "{}", e, b.st.s);
    });
}

int mael_dense_restore(mael_dense* d, const char* path) {
    if (!d || !path) return MAEL_E_ARG;
    return guard(d->ctx, [&] {
        mael::DenseBasin& b = *d->b;
        Reader r(path);
        if (r.kind != 0) throw mael::Error(mael::E_ARG, "not a dense checkpoint");
        hipStream_t s = b.st.s;
        b.st.sync();
        const int64_t C = (int64_t)b.cells();
        r.upload("phi", b.phi.p, C, 4, s, C);
        r.upload("types", b.types.p, C, 1, s, C);
        r.upload("P", b.P.p, C, 4, s, C);
        const char* un[3] = {"u", "v", "w"};
        const char* fn[3] = {"fu", "fv", "fw"};
        for (int a = 0; a < 3; ++a) {
            r.upload(un[a], b.u[a].p, (int64_t)b.faces(a), 4, s, (int64_t)b.faces(a));
            r.upload(fn[a], b.flag[a].p, (int64_t)b.faces(a), 1, s, (int64_t)b.faces(a));
            MAEL_HIP(hipMemcpyAsync(b.u0[a].p, b.u[a].p, b.faces(a) * sizeof(float), hipMemcpyDeviceToDevice, s));
        }
        MAEL_HIP(hipMemcpyAsync(b.phi0.p, b.phi.p, C * sizeof(float), hipMemcpyDeviceToDevice, s));
        const int64_t n = (int64_t)r.find("px").count;
        if (n > b.capacity) throw mael::Error(mael::E_ARG, "Particle capacity is too small");
        for (int f = 0; f < mael::PFIELDS; ++f) r.upload(PNAMES[f], b.pbuf[b.cur][f].p, b.capacity, 4, s, n);
        b.n = n;
        b.st.sync();
    });
}

int mael_sparse_checkpoint(mael_sparse* sp, const char* path, const char* meta) {
    if (!sp || !path) return MAEL_E_ARG;
    return guard(sp->ctx, [&] {
        mael::SparseBasin& b = *sp->b;
        b.st.sync();
        const int64_t ns = b.nslots, Cc = (int64_t)b.ncells();
        const int64_t fc = ns * mael::SBCELLS, ff = ns * mael::SBFACES;
        std::vector<Entry> e = {{"phi_f", b.phi.fine.p, fc, 4},     {"phi_c", b.phi.coarse.p, Cc, 4},
                                {"types_f", b.types.fine.p, fc, 1}, {"types_c", b.types.coarse.p, Cc, 1},
                                {"P_f", b.P.fine.p, fc, 4},         {"P_c", b.P.coarse.p, Cc, 4}};
        const char* un[3] = {"u", "v", "w"};
        const char* fn[3] = {"fu", "fv", "fw"};
        for (int a = 0; a < 3; ++a) {
            e.push_back({std::string(un[a]) + "_f", b.u[a].fine.p, ff, 4});
            e.push_back({std::string(un[a]) + "_c", b.u[a].coarse.p, (int64_t)b.nfaces(a), 4});
            e.push_back({std::string(fn[a]) + "_f", b.fl[a].fine.p, ff, 1});
            e.push_back({std::string(fn[a]) + "_c", b.fl[a].coarse.p, (int64_t)b.nfaces(a), 1});
        }
        e.push_back({"slot", b.slot.p, (int64_t)b.nblocks(), 4});
        // the host's topology: the live slots' blocks and the free list
        std::vector<int32_t> where(3 * ns), freel(b.freel.begin(), b.freel.end());
        for (int64_t k = 0; k < ns; ++k) {
            where[3 * k] = b.hwhere[k].x; where[3 * k + 1] = b.hwhere[k].y; where[3 * k + 2] = b.hwhere[k].z;
        }
        e.push_back({"where", nullptr, (int64_t)where.size(), 4, where.data()});
        e.push_back({"free", nullptr, (int64_t)freel.size(), 4, freel.data()});
        for (int f = 0; f < mael::PFIELDS; ++f) e.push_back({PNAMES[f], b.pbuf[b.cur][f].p, b.n, 4});
        write_checkpoint(path, 1, meta ? meta : "{}", e, b.st.s);
    });
}

int mael_sparse_restore(mael_sparse* sp, const char* path) {
    if (!sp || !path) return MAEL_E_ARG;
    int st = guard(sp->ctx, [&] {
        mael::SparseBasin& b = *sp->b;
        Reader r(path);
        if (r.kind != 1) throw mael::Error(mael::E_ARG, "not a sparse checkpoint");
        std::vector<int32_t> slot = r.host<int32_t>("slot"), where3 = r.host<int32_t>("where"),
                             freel = r.host<int32_t>("free");
        if (slot.size() != b.nblocks()) throw mael::Error(mael::E_ARG, "Checkpoint grid differs");
        const int32_t ns = (int32_t)(where3.size() / 3);
        if (ns > b.blocks) throw mael::Error(mael::E_ARG, "the checkpoint holds more blocks than this basin's pool");
        std::vector<int32_t> where(3 * (size_t)b.blocks, -1);
        std::copy(where3.begin(), where3.end(), where.begin());
        int tst = mael_sparse_set_topology(sp, slot.data(), where.data(), ns, freel.data(), (int32_t)freel.size());
        if (tst != MAEL_OK) throw mael::Error(tst, sp->ctx->err);
        hipStream_t s = b.st.s;
        const int64_t Cc = (int64_t)b.ncells(), fc = (int64_t)ns * mael::SBCELLS, ff = (int64_t)ns * mael::SBFACES;
        const int64_t capc = (int64_t)b.blocks * mael::SBCELLS, capf = (int64_t)b.blocks * mael::SBFACES;
        r.upload("phi_f", b.phi.fine.p, capc, 4, s, fc);
        r.upload("phi_c", b.phi.coarse.p, Cc, 4, s, Cc);
        r.upload("types_f", b.types.fine.p, capc, 1, s, fc);
        r.upload("types_c", b.types.coarse.p, Cc, 1, s, Cc);
        r.upload("P_f", b.P.fine.p, capc, 4, s, fc);
        r.upload("P_c", b.P.coarse.p, Cc, 4, s, Cc);
        const char* un[3] = {"u", "v", "w"};
        const char* fn[3] = {"fu", "fv", "fw"};
        for (int a = 0; a < 3; ++a) {
            const int64_t Fa = (int64_t)b.nfaces(a);
            r.upload(std::string(un[a]) + "_f", b.u[a].fine.p, capf, 4, s, ff);
            r.upload(std::string(un[a]) + "_c", b.u[a].coarse.p, Fa, 4, s, Fa);
            r.upload(std::string(fn[a]) + "_f", b.fl[a].fine.p, capf, 1, s, ff);
            r.upload(std::string(fn[a]) + "_c", b.fl[a].coarse.p, Fa, 1, s, Fa);
        }
        const int64_t n = (int64_t)r.find("px").count;
        if (n > b.capacity) throw mael::Error(mael::E_ARG, "Particle capacity is too small");
        for (int f = 0; f < mael::PFIELDS; ++f) r.upload(PNAMES[f], b.pbuf[b.cur][f].p, b.capacity, 4, s, n);
        b.n = n;
        b.keep();   // φ0 = φ, u0 = u (checkpoint.jl:54-57)
        b.st.sync();
    });
    return st;
}

}  // extern "C"
