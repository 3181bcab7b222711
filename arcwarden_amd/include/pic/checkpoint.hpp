// pic/checkpoint.hpp -- checkpoint / restart files (ARCW v1).
//
// Reference: include/pic/checkpoint.hpp:1-131.  CheckpointHeader,
// CheckpointArray, checkpoint_write_header and checkpoint_read_header keep
// the reference's names and behaviour (little-endian header, deck and git
// blocks, array manifest; "checkpoint: bad magic" and "checkpoint: version N
// newer than reader" as std::runtime_error).  Beyond the reference, the
// payloads are written and read too: checkpoint_write takes each array's
// data (host or device memory) and CheckpointReader::read fills a buffer;
// the file I/O runs in the library (aw_checkpoint_*).
#ifndef ARC_PIC_CHECKPOINT_HPP
#define ARC_PIC_CHECKPOINT_HPP

#include "pic/hip_utils.hpp"

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace arc {

struct CheckpointArray {
    std::string name;
    uint32_t dtype = 0;  // 0 f32, 1 f64, 2 i32, 3 u64
    uint64_t count = 0;
};

struct CheckpointHeader {
    static constexpr uint32_t MAGIC = 0x57435241u;  // "ARCW"
    static constexpr uint32_t VERSION = 1;

    uint64_t rng_seed = 0;
    int64_t step = 0;
    double time = 0.0;
    double sd = 1.0;
    double eps_l = 0.0;
    std::string deck;
    std::string git_hash;
    std::vector<CheckpointArray> manifest;
};

// One array's payload for checkpoint_write: host or device memory.
struct CheckpointPayload {
    const void* data = nullptr;
    bool on_device = false;
};

namespace detail {
inline void ck_check(aw_status st) {
    if (st != AW_OK) throw std::runtime_error(last_error());
}
}  // namespace detail

// Header, deck, git hash and manifest; with payloads (one per manifest entry,
// in order) the arrays follow.
inline void checkpoint_write(const std::string& path, const CheckpointHeader& h,
                             const std::vector<CheckpointPayload>& payloads = {}) {
    if (!payloads.empty() && payloads.size() != h.manifest.size())
        throw std::invalid_argument("checkpoint_write: one payload per manifest entry");
    aw_checkpoint_header c{};
    c.rng_seed = h.rng_seed;
    c.step = h.step;
    c.time = h.time;
    c.sd = h.sd;
    c.eps_l = h.eps_l;
    std::vector<aw_checkpoint_array> arr;
    for (std::size_t i = 0; i < h.manifest.size(); ++i) {
        aw_checkpoint_array a{};
        a.name = h.manifest[i].name.c_str();
        a.dtype = h.manifest[i].dtype;
        a.count = h.manifest[i].count;
        if (!payloads.empty()) {
            a.data = const_cast<void*>(payloads[i].data);
            a.on_device = payloads[i].on_device ? 1 : 0;
        }
        arr.push_back(a);
    }
    detail::ck_check(aw_checkpoint_write(path.c_str(), &c, h.deck.c_str(), h.git_hash.c_str(),
                                         arr.data(), static_cast<int32_t>(arr.size())));
}

inline void checkpoint_write_header(const std::string& path, const CheckpointHeader& h) {
    checkpoint_write(path, h);
}

// An opened checkpoint: its header and the payloads on demand.
class CheckpointReader {
public:
    explicit CheckpointReader(const std::string& path) {
        detail::ck_check(aw_checkpoint_open(path.c_str(), &ck_));
        aw_checkpoint_header c{};
        int32_t na = 0;
        int64_t dl = 0, gl = 0;
        detail::ck_check(aw_checkpoint_info(ck_, &c, &na, &dl, &gl));
        h_.rng_seed = c.rng_seed;
        h_.step = c.step;
        h_.time = c.time;
        h_.sd = c.sd;
        h_.eps_l = c.eps_l;
        h_.deck = text(0, dl);
        h_.git_hash = text(1, gl);
        for (int32_t i = 0; i < na; ++i) {
            char name[4097];
            CheckpointArray a;
            int32_t present = 0;
            detail::ck_check(aw_checkpoint_array_info(ck_, i, name, sizeof name, &a.dtype,
                                                      &a.count, &present));
            a.name = name;
            present_.push_back(present != 0);
            h_.manifest.push_back(a);
        }
    }
    ~CheckpointReader() {
        if (ck_) (void)aw_checkpoint_close(ck_);
    }
    CheckpointReader(const CheckpointReader&) = delete;
    CheckpointReader& operator=(const CheckpointReader&) = delete;

    const CheckpointHeader& header() const { return h_; }
    bool has_payload(std::size_t i) const { return present_.at(i); }
    void read(std::size_t i, void* dst, bool on_device = false) const {
        detail::ck_check(
            aw_checkpoint_read_array(ck_, static_cast<int32_t>(i), dst, on_device ? 1 : 0));
    }

private:
    std::string text(int which, int64_t len) {
        std::string s(static_cast<std::size_t>(len) + 1, '\0');
        detail::ck_check(aw_checkpoint_text(ck_, which, s.data(), len + 1));
        s.resize(static_cast<std::size_t>(len));
        return s;
    }
    aw_checkpoint* ck_ = nullptr;
    CheckpointHeader h_;
    std::vector<bool> present_;
};

inline CheckpointHeader checkpoint_read_header(const std::string& path) {
    return CheckpointReader(path).header();
}

}  // namespace arc

#endif  // ARC_PIC_CHECKPOINT_HPP
