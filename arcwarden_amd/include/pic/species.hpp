// pic/species.hpp -- host-side description of one particle species.
//
// Reference: Species / SpeciesList (include/pic/species.hpp:20-47).
#ifndef ARC_PIC_SPECIES_HPP
#define ARC_PIC_SPECIES_HPP

#include "aw.h"

#include <cstdio>
#include <string>
#include <vector>

namespace arc {

struct Species {
    std::string name = "electrons";
    double density = 1.0;
    int ppc = 0;
    double uth[3] = {0.0, 0.0, 0.0};
    double ufl[3] = {0.0, 0.0, 0.0};
    bool deltaf = false;
    int dist = 0;
    double lc_rho = 1.0;
    double lc_kappa = 0.3;
    double taud = 0.0;
    double wdnoise = 0.0;

    aw_species to_c() const {
        aw_species s;
        aw_species_defaults(&s);
        std::snprintf(s.name, sizeof(s.name), "%s", name.c_str());
        s.density = density;
        s.ppc = ppc;
        for (int i = 0; i < 3; ++i) {
            s.uth[i] = uth[i];
            s.ufl[i] = ufl[i];
        }
        s.deltaf = deltaf ? 1 : 0;
        s.dist = dist;
        s.lc_rho = lc_rho;
        s.lc_kappa = lc_kappa;
        s.taud = taud;
        s.wdnoise = wdnoise;
        return s;
    }

    static Species from_c(const aw_species& c) {
        Species s;
        s.name = c.name;
        s.density = c.density;
        s.ppc = c.ppc;
        for (int i = 0; i < 3; ++i) {
            s.uth[i] = c.uth[i];
            s.ufl[i] = c.ufl[i];
        }
        s.deltaf = c.deltaf != 0;
        s.dist = c.dist;
        s.lc_rho = c.lc_rho;
        s.lc_kappa = c.lc_kappa;
        s.taud = c.taud;
        s.wdnoise = c.wdnoise;
        return s;
    }
};

using SpeciesList = std::vector<Species>;

inline std::vector<aw_species> to_c(const SpeciesList& list) {
    std::vector<aw_species> out;
    out.reserve(list.size());
    for (const auto& s : list) out.push_back(s.to_c());
    return out;
}

}  // namespace arc

#endif  // ARC_PIC_SPECIES_HPP
