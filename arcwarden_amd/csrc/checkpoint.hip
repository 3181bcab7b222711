// checkpoint.hip -- checkpoint / restart files in the ARCW v1 layout with
// their array payloads.
//
// Reference: include/pic/checkpoint.hpp:1-131 (the layout and the header
// reader / writer; the reference defines the payload section but streams no
// arrays).  Layout, little-endian, one file:
//   header    "ARCW" (u32 0x57435241), u32 version = 1, u64 rng_seed, i64 step,
//             f64 time, f64 S_d, f64 eps_L
//   deck      u64 nbytes + the deck text
//   git       u64 nbytes + the git hash
//   manifest  u32 narrays, then per array u32 name_len + name, u32 dtype
//             (0 f32, 1 f64, 2 i32, 3 u64), u64 count
//   payloads  the arrays' raw bytes in manifest order
// Readers refuse another magic and a newer version; a file whose payloads
// are absent (schema only) reads as a header.  Payloads may live on the host
// or the device: device arrays are staged through a pinned host buffer in
// 64 MiB pieces, so a 16.7 M-particle run (335 MB of particle data) never
// needs a host copy of its whole state.
#include "aw_internal.hpp"

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace aw {
namespace {

constexpr uint32_t kMagic = 0x57435241u;  // "ARCW"
constexpr uint32_t kVersion = 1;
constexpr size_t kStage = 64ull << 20;

size_t dtype_size(uint32_t t) {
    switch (t) {
        case AW_CK_F32: return 4;
        case AW_CK_F64: return 8;
        case AW_CK_I32: return 4;
        case AW_CK_U64: return 8;
    }
    fail(AW_ERR_INVALID, "checkpoint: unknown dtype " + std::to_string(t), __FILE__, __LINE__);
    return 0;
}

struct File {
    std::FILE* f = nullptr;
    ~File() {
        if (f) std::fclose(f);
    }
};

void put(std::FILE* f, const void* p, size_t n) {
    if (n && std::fwrite(p, 1, n, f) != n)
        fail(AW_ERR_INVALID, "checkpoint: short write", __FILE__, __LINE__);
}

void get(std::FILE* f, void* p, size_t n) {
    if (n && std::fread(p, 1, n, f) != n)
        fail(AW_ERR_INVALID, "checkpoint: short read", __FILE__, __LINE__);
}

void put_str(std::FILE* f, const char* s) {
    const uint64_t n = s ? std::strlen(s) : 0;
    put(f, &n, 8);
    put(f, s, n);
}

std::string get_str(std::FILE* f) {
    uint64_t n = 0;
    get(f, &n, 8);
    if (n > (1ull << 30)) fail(AW_ERR_INVALID, "checkpoint: absurd string length", __FILE__, __LINE__);
    std::string s(n, '\0');
    get(f, s.data(), n);
    return s;
}

// Copies bytes between a file and host or device memory in staged pieces.
struct Stager {
    void* pinned = nullptr;
    ~Stager() {
        if (pinned) (void)hipHostFree(pinned);
    }
    void* buf() {
        if (!pinned) AW_HIP(hipHostMalloc(&pinned, kStage, hipHostMallocDefault));
        return pinned;
    }
    void write(std::FILE* f, const void* src, size_t n, bool device) {
        if (!device) {
            put(f, src, n);
            return;
        }
        const char* s = static_cast<const char*>(src);
        for (size_t o = 0; o < n; o += kStage) {
            const size_t k = std::min(kStage, n - o);
            AW_HIP(hipMemcpy(buf(), s + o, k, hipMemcpyDeviceToHost));
            put(f, buf(), k);
        }
    }
    void read(std::FILE* f, void* dst, size_t n, bool device) {
        if (!device) {
            get(f, dst, n);
            return;
        }
        char* d = static_cast<char*>(dst);
        for (size_t o = 0; o < n; o += kStage) {
            const size_t k = std::min(kStage, n - o);
            get(f, buf(), k);
            AW_HIP(hipMemcpy(d + o, buf(), k, hipMemcpyHostToDevice));
        }
    }
};

}  // namespace
}  // namespace aw

using namespace aw;

// An opened checkpoint file: the parsed header and the payload offsets.
struct aw_checkpoint {
    aw_checkpoint_header h{};
    std::string deck, git;
    struct Entry {
        std::string name;
        uint32_t dtype;
        uint64_t count;
        long offset;  // of the payload, -1 when the file ends before it
    };
    std::vector<Entry> arrays;
    std::string path;
};

extern "C" {

aw_status aw_checkpoint_write(const char* path, const aw_checkpoint_header* h, const char* deck,
                              const char* git_hash, const aw_checkpoint_array* arrays,
                              int32_t narrays) {
    return guard("aw_checkpoint_write", [&] {
        AW_REQUIRE(path && h, "NULL argument");
        AW_REQUIRE(narrays >= 0 && (narrays == 0 || arrays), "arrays is NULL");
        File file;
        file.f = std::fopen(path, "wb");
        if (!file.f) fail(AW_ERR_INVALID, std::string("checkpoint: cannot open ") + path, __FILE__, __LINE__);
        std::FILE* f = file.f;
        put(f, &kMagic, 4);
        put(f, &kVersion, 4);
        put(f, &h->rng_seed, 8);
        put(f, &h->step, 8);
        put(f, &h->time, 8);
        put(f, &h->sd, 8);
        put(f, &h->eps_l, 8);
        put_str(f, deck);
        put_str(f, git_hash);
        const uint32_t na = (uint32_t)narrays;
        put(f, &na, 4);
        for (int32_t i = 0; i < narrays; ++i) {
            const aw_checkpoint_array& a = arrays[i];
            AW_REQUIRE(a.name, "array name is NULL");
            (void)dtype_size(a.dtype);
            const uint32_t nl = (uint32_t)std::strlen(a.name);
            put(f, &nl, 4);
            put(f, a.name, nl);
            put(f, &a.dtype, 4);
            put(f, &a.count, 8);
        }
        Stager st;
        for (int32_t i = 0; i < narrays; ++i) {
            const aw_checkpoint_array& a = arrays[i];
            if (!a.data) continue;  // schema-only entry: the payloads stop here
            st.write(f, a.data, (size_t)a.count * dtype_size(a.dtype), a.on_device != 0);
        }
        if (std::fflush(f) != 0) fail(AW_ERR_INVALID, "checkpoint: flush failed", __FILE__, __LINE__);
    });
}

aw_status aw_checkpoint_open(const char* path, aw_checkpoint** out) {
    return guard("aw_checkpoint_open", [&] {
        AW_REQUIRE(path && out, "NULL argument");
        *out = nullptr;
        File file;
        file.f = std::fopen(path, "rb");
        if (!file.f) fail(AW_ERR_INVALID, std::string("checkpoint: cannot open ") + path, __FILE__, __LINE__);
        std::FILE* f = file.f;
        auto ck = std::make_unique<aw_checkpoint>();
        ck->path = path;
        uint32_t magic = 0, ver = 0;
        get(f, &magic, 4);
        get(f, &ver, 4);
        if (magic != kMagic) fail(AW_ERR_INVALID, "checkpoint: bad magic", __FILE__, __LINE__);
        if (ver > kVersion)
            fail(AW_ERR_INVALID,
                 "checkpoint: version " + std::to_string(ver) + " newer than reader", __FILE__,
                 __LINE__);
        ck->h.version = ver;
        get(f, &ck->h.rng_seed, 8);
        get(f, &ck->h.step, 8);
        get(f, &ck->h.time, 8);
        get(f, &ck->h.sd, 8);
        get(f, &ck->h.eps_l, 8);
        ck->deck = get_str(f);
        ck->git = get_str(f);
        uint32_t na = 0;
        get(f, &na, 4);
        if (na > (1u << 20)) fail(AW_ERR_INVALID, "checkpoint: absurd array count", __FILE__, __LINE__);
        for (uint32_t i = 0; i < na; ++i) {
            aw_checkpoint::Entry e;
            uint32_t nl = 0;
            get(f, &nl, 4);
            if (nl > 4096) fail(AW_ERR_INVALID, "checkpoint: absurd name length", __FILE__, __LINE__);
            e.name.resize(nl);
            get(f, e.name.data(), nl);
            get(f, &e.dtype, 4);
            get(f, &e.count, 8);
            (void)dtype_size(e.dtype);
            ck->arrays.push_back(std::move(e));
        }
        // payload offsets; entries beyond the end of the file are absent
        long off = std::ftell(f);
        std::fseek(f, 0, SEEK_END);
        const long end = std::ftell(f);
        for (auto& e : ck->arrays) {
            const long n = (long)(e.count * dtype_size(e.dtype));
            if (off >= 0 && off + n <= end) {
                e.offset = off;
                off += n;
            } else {
                e.offset = -1;
                off = -1;
            }
        }
        *out = ck.release();
    });
}

aw_status aw_checkpoint_close(aw_checkpoint* ck) {
    return guard("aw_checkpoint_close", [&] { delete ck; });
}

aw_status aw_checkpoint_info(const aw_checkpoint* ck, aw_checkpoint_header* h, int32_t* narrays,
                             int64_t* deck_len, int64_t* git_len) {
    return guard("aw_checkpoint_info", [&] {
        AW_REQUIRE(ck, "checkpoint is NULL");
        if (h) *h = ck->h;
        if (narrays) *narrays = (int32_t)ck->arrays.size();
        if (deck_len) *deck_len = (int64_t)ck->deck.size();
        if (git_len) *git_len = (int64_t)ck->git.size();
    });
}

aw_status aw_checkpoint_text(const aw_checkpoint* ck, int32_t which, char* buf, int64_t cap) {
    return guard("aw_checkpoint_text", [&] {
        AW_REQUIRE(ck && buf, "NULL argument");
        const std::string& s = which == 0 ? ck->deck : ck->git;
        AW_REQUIRE(cap > (int64_t)s.size(), "buffer too small (needs length + 1)");
        std::memcpy(buf, s.data(), s.size());
        buf[s.size()] = 0;
    });
}

aw_status aw_checkpoint_array_info(const aw_checkpoint* ck, int32_t i, char* name, int64_t cap,
                                   uint32_t* dtype, uint64_t* count, int32_t* present) {
    return guard("aw_checkpoint_array_info", [&] {
        AW_REQUIRE(ck, "checkpoint is NULL");
        AW_REQUIRE(i >= 0 && i < (int32_t)ck->arrays.size(), "array index out of range");
        const auto& e = ck->arrays[(size_t)i];
        if (name) {
            AW_REQUIRE(cap > (int64_t)e.name.size(), "name buffer too small");
            std::memcpy(name, e.name.data(), e.name.size());
            name[e.name.size()] = 0;
        }
        if (dtype) *dtype = e.dtype;
        if (count) *count = e.count;
        if (present) *present = e.offset >= 0 ? 1 : 0;
    });
}

aw_status aw_checkpoint_read_array(const aw_checkpoint* ck, int32_t i, void* dst,
                                   int32_t on_device) {
    return guard("aw_checkpoint_read_array", [&] {
        AW_REQUIRE(ck && dst, "NULL argument");
        AW_REQUIRE(i >= 0 && i < (int32_t)ck->arrays.size(), "array index out of range");
        const auto& e = ck->arrays[(size_t)i];
        if (e.offset < 0)
            fail(AW_ERR_INVALID, "checkpoint: payload of '" + e.name + "' is absent", __FILE__, __LINE__);
        File file;
        file.f = std::fopen(ck->path.c_str(), "rb");
        if (!file.f) fail(AW_ERR_INVALID, "checkpoint: cannot reopen " + ck->path, __FILE__, __LINE__);
        std::fseek(file.f, e.offset, SEEK_SET);
        Stager st;
        st.read(file.f, dst, (size_t)e.count * dtype_size(e.dtype), on_device != 0);
    });
}

}  // extern "C"
