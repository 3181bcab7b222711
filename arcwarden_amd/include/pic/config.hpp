// pic/config.hpp -- compile-time SimConfig and run-time RunParams.
//
// Reference: include/pic/config.hpp:46-313.  SimConfig keeps the reference's
// template parameters and policy tags; the tags select the library's deposit
// policy (AW_DEPOSIT_*) and field model (AW_MODEL_*) at compile time.
// RunParams is the C ABI's aw_run_params with the reference's member
// initialisers (the normalisation closure eps0 = me = |e| = 1, wpe = 1).
#ifndef ARC_PIC_CONFIG_HPP
#define ARC_PIC_CONFIG_HPP

#include "aw.h"

#include <type_traits>

namespace arc {

enum class ShapeOrder { CIC, TSC };

struct AtomicGlobalDeposit {};  // one particle per thread, global atomics
struct SharedTileDeposit {};    // whole grid privatised in LDS (small grids)
struct CellOwnedDeposit {};     // the reference's per-cell variant: runs as tiled+binned
struct TiledBinnedDeposit {};   // spatial tiles + binning, LDS tile + apron

enum class FieldModel { Electrostatic, Darwin };

template <class DepositPolicy>
inline constexpr int deposit_policy_id() {
    if constexpr (std::is_same_v<DepositPolicy, AtomicGlobalDeposit>) return AW_DEPOSIT_GLOBAL;
    else if constexpr (std::is_same_v<DepositPolicy, SharedTileDeposit>) return AW_DEPOSIT_LDS_GRID;
    else return AW_DEPOSIT_BINNED;
}

template <int Dim, int VelDim, ShapeOrder Shape, typename Real, bool HasB0,
          class DepositPolicy = AtomicGlobalDeposit,
          FieldModel Model = FieldModel::Electrostatic>
struct SimConfig {
    static constexpr int dim = Dim;
    static constexpr int vdim = VelDim;
    static constexpr ShapeOrder shape = Shape;
    static constexpr bool has_b0 = HasB0;
    static constexpr FieldModel field_model = Model;
    using real = Real;
    using deposit = DepositPolicy;
    static constexpr int aw_deposit = deposit_policy_id<DepositPolicy>();
    static constexpr int aw_model = Model == FieldModel::Darwin ? AW_MODEL_DARWIN : AW_MODEL_ES;

    static_assert(Dim == 2 && VelDim == 3, "this library implements the 2D3V path");
    static_assert(Shape == ShapeOrder::CIC, "this library implements CIC");
    static_assert(std::is_same_v<Real, float>, "particles and fields are FP32");
};

using Cfg = SimConfig<2, 3, ShapeOrder::CIC, float, true, SharedTileDeposit>;
using CfgDarwin =
    SimConfig<2, 3, ShapeOrder::CIC, float, true, TiledBinnedDeposit, FieldModel::Darwin>;

enum class NormMode { OmegaPeUnity };

// RunParams: member for member the reference's (bool members as int32_t).
struct RunParams : aw_run_params {
    RunParams() : aw_run_params{} { aw_run_params_defaults(this); }
};

}  // namespace arc

#endif  // ARC_PIC_CONFIG_HPP
