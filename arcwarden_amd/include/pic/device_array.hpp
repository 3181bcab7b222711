// pic/device_array.hpp -- device memory owners and views.
//
// Reference: DeviceView<T> / DeviceArray<T> (include/pic/device_array.hpp:25-107).
// DeviceArray is move-only and either owns its memory (allocated through the
// library, aw_device_alloc) or borrows it (arrays owned by a Simulation or
// by a contiguous block); data(), size(), bytes(), zero() behave as the
// reference's.
#ifndef ARC_PIC_DEVICE_ARRAY_HPP
#define ARC_PIC_DEVICE_ARRAY_HPP

#include "pic/hip_utils.hpp"

#include <cstddef>
#include <utility>
#include <vector>

namespace arc {

template <class T>
struct DeviceView {
    T* ptr = nullptr;
    int n = 0;
    int size() const { return n; }
    T* data() const { return ptr; }
};

template <class T>
class DeviceArray {
public:
    DeviceArray() = default;
    explicit DeviceArray(std::size_t n) : n_(n), owned_(true) {
        void* p = nullptr;
        check(aw_device_alloc(n * sizeof(T), &p));
        ptr_ = static_cast<T*>(p);
    }
    // Non-owning array over memory that lives elsewhere.
    static DeviceArray borrow(T* p, std::size_t n) {
        DeviceArray a;
        a.ptr_ = p;
        a.n_ = n;
        return a;
    }
    ~DeviceArray() { reset(); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    DeviceArray(DeviceArray&& o) noexcept { swap(o); }
    DeviceArray& operator=(DeviceArray&& o) noexcept {
        if (this != &o) {
            reset();
            swap(o);
        }
        return *this;
    }

    T* data() { return ptr_; }
    const T* data() const { return ptr_; }
    std::size_t size() const { return n_; }
    std::size_t bytes() const { return n_ * sizeof(T); }
    bool owned() const { return owned_; }
    DeviceView<T> view() const { return DeviceView<T>{ptr_, static_cast<int>(n_)}; }

    void zero(hipStream_t stream = nullptr) {
        if (n_) check(aw_memset_zero(ptr_, bytes(), to_aw(stream)));
    }
    std::vector<T> to_host() const {
        std::vector<T> h(n_);
        if (n_) check(aw_copy(h.data(), ptr_, bytes()));
        return h;
    }
    void from_host(const std::vector<T>& h) {
        if (h.size() != n_) throw std::length_error("DeviceArray::from_host: size mismatch");
        if (n_) check(aw_copy(ptr_, h.data(), bytes()));
    }
    void reset() noexcept {
        if (owned_ && ptr_) (void)aw_device_free(ptr_);
        ptr_ = nullptr;
        n_ = 0;
        owned_ = false;
    }

private:
    void swap(DeviceArray& o) noexcept {
        std::swap(ptr_, o.ptr_);
        std::swap(n_, o.n_);
        std::swap(owned_, o.owned_);
    }
    T* ptr_ = nullptr;
    std::size_t n_ = 0;
    bool owned_ = false;
};

}  // namespace arc

#endif  // ARC_PIC_DEVICE_ARRAY_HPP
