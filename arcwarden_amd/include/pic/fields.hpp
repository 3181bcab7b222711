// pic/fields.hpp -- real-space field grids.
//
// Reference: Fields (include/pic/fields.hpp:24-73).  allocate() makes Ex, Ey;
// allocate_em() adds Ez, B and the Darwin split E_L / E_T, laid out as the
// library's solves write them: [ELx|ELy|Bx|By|Bz] and [ETx|ETy|ETz] blocks.
#ifndef ARC_PIC_FIELDS_HPP
#define ARC_PIC_FIELDS_HPP

#include "pic/device_array.hpp"
#include "pic/grid.hpp"

namespace arc {

struct Fields {
    DeviceArray<float> Ex, Ey;
    DeviceArray<float> Ez, Bx, By, Bz;
    DeviceArray<float> ELx, ELy;
    DeviceArray<float> ETx, ETy, ETz;

    void allocate(const Grid& g) {
        const std::size_t n = static_cast<std::size_t>(g.real_size());
        Ex = DeviceArray<float>(n);
        Ey = DeviceArray<float>(n);
    }

    void allocate_em(const Grid& g, hipStream_t stream = nullptr) {
        const std::size_t n = static_cast<std::size_t>(g.real_size());
        if (Ex.size() != n) allocate(g);
        Ez = DeviceArray<float>(n);
        elb_ = DeviceArray<float>(5 * n);
        et_ = DeviceArray<float>(3 * n);
        float* a = elb_.data();
        float* b = et_.data();
        ELx = DeviceArray<float>::borrow(a, n);
        ELy = DeviceArray<float>::borrow(a + n, n);
        Bx = DeviceArray<float>::borrow(a + 2 * n, n);
        By = DeviceArray<float>::borrow(a + 3 * n, n);
        Bz = DeviceArray<float>::borrow(a + 4 * n, n);
        ETx = DeviceArray<float>::borrow(b, n);
        ETy = DeviceArray<float>::borrow(b + n, n);
        ETz = DeviceArray<float>::borrow(b + 2 * n, n);
        Ez.zero(stream);
        elb_.zero(stream);
        et_.zero(stream);  // the retained E_T starts at zero
    }

    void zero(hipStream_t stream = nullptr) {
        Ex.zero(stream);
        Ey.zero(stream);
    }

    float* elb() { return elb_.data(); }  // [ELx|ELy|Bx|By|Bz]
    float* et() { return et_.data(); }    // [ETx|ETy|ETz]

    aw_fields c_view() {
        aw_fields f{};
        f.Ex = Ex.data();
        f.Ey = Ey.data();
        f.Ez = Ez.data();
        f.Bx = Bx.data();
        f.By = By.data();
        f.Bz = Bz.data();
        f.ELx = ELx.data();
        f.ELy = ELy.data();
        f.ETx = ETx.data();
        f.ETy = ETy.data();
        f.ETz = ETz.data();
        return f;
    }

private:
    DeviceArray<float> elb_, et_;
};

}  // namespace arc

#endif  // ARC_PIC_FIELDS_HPP
