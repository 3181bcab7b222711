// pic/sources.hpp -- charge / current source grids.
//
// Reference: Sources (include/pic/sources.hpp:23-85).  Every grid is FP32
// [ny][nx].  allocate_em() lays rho|Jx|Jy|Jz and dcux|dcuy|dcuz|amu0..amu3 out
// as two contiguous blocks (the Darwin solves read them as batches); the
// members are views into those blocks.
#ifndef ARC_PIC_SOURCES_HPP
#define ARC_PIC_SOURCES_HPP

#include "pic/device_array.hpp"
#include "pic/grid.hpp"

namespace arc {

struct Sources {
    DeviceArray<float> rho;
    DeviceArray<float> Jx, Jy, Jz;
    DeviceArray<float> dcux, dcuy, dcuz;
    DeviceArray<float> amu0, amu1, amu2, amu3;

    void allocate(const Grid& g) {
        block4_.reset();
        rho = DeviceArray<float>(static_cast<std::size_t>(g.real_size()));
    }

    void allocate_em(const Grid& g) {
        const std::size_t n = static_cast<std::size_t>(g.real_size());
        block4_ = DeviceArray<float>(4 * n);
        block7_ = DeviceArray<float>(7 * n);
        float* a = block4_.data();
        float* b = block7_.data();
        rho = DeviceArray<float>::borrow(a, n);
        Jx = DeviceArray<float>::borrow(a + n, n);
        Jy = DeviceArray<float>::borrow(a + 2 * n, n);
        Jz = DeviceArray<float>::borrow(a + 3 * n, n);
        dcux = DeviceArray<float>::borrow(b, n);
        dcuy = DeviceArray<float>::borrow(b + n, n);
        dcuz = DeviceArray<float>::borrow(b + 2 * n, n);
        amu0 = DeviceArray<float>::borrow(b + 3 * n, n);
        amu1 = DeviceArray<float>::borrow(b + 4 * n, n);
        amu2 = DeviceArray<float>::borrow(b + 5 * n, n);
        amu3 = DeviceArray<float>::borrow(b + 6 * n, n);
    }

    bool em() const { return block4_.size() != 0; }

    void zero(hipStream_t stream = nullptr) { rho.zero(stream); }
    void zero_rho_j(hipStream_t stream = nullptr) {
        if (em()) block4_.zero(stream);
        else rho.zero(stream);
    }
    void zero_dcu_amu(hipStream_t stream = nullptr) { block7_.zero(stream); }
    void zero_dcu(hipStream_t stream = nullptr) {
        dcux.zero(stream);
        dcuy.zero(stream);
        dcuz.zero(stream);
    }

    // [rho|Jx|Jy|Jz] and [dcux|dcuy|dcuz|amu0..amu3] for the Darwin solves
    const float* src4() const { return block4_.data(); }
    const float* src7() const { return block7_.data(); }

    aw_sources c_view() {
        aw_sources s{};
        s.rho = rho.data();
        s.Jx = Jx.data();
        s.Jy = Jy.data();
        s.Jz = Jz.data();
        s.dcux = dcux.data();
        s.dcuy = dcuy.data();
        s.dcuz = dcuz.data();
        s.amu0 = amu0.data();
        s.amu1 = amu1.data();
        s.amu2 = amu2.data();
        s.amu3 = amu3.data();
        return s;
    }

private:
    DeviceArray<float> block4_, block7_;
};

}  // namespace arc

#endif  // ARC_PIC_SOURCES_HPP
