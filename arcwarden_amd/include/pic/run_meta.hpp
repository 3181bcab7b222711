// pic/run_meta.hpp -- provenance of a deck-driven run.
//
// Reference: write_run_meta / git_hash_best_effort
// (include/pic/run_meta.hpp:17-55).  The output directory receives
// deck_snapshot.ini (a byte copy of the deck) and RUN_META.txt with the keys
// timestamp, git_hash, deck, S_d, eps_L and cmdline.
#ifndef ARC_PIC_RUN_META_HPP
#define ARC_PIC_RUN_META_HPP

#include <cstdio>
#include <ctime>
#include <fstream>
#include <string>

namespace arc {

// HEAD of the git repository around the working directory, or "unknown".
inline std::string git_hash_best_effort() {
    std::string hash = "unknown";
    if (std::FILE* p = ::popen("git rev-parse HEAD 2>/dev/null", "r")) {
        char buf[96] = {0};
        if (std::fgets(buf, sizeof buf, p)) {
            std::string s(buf);
            while (!s.empty() && (s.back() == '\n' || s.back() == '\r')) s.pop_back();
            if (!s.empty()) hash = s;
        }
        ::pclose(p);
    }
    return hash;
}

inline void write_run_meta(const std::string& outdir, const std::string& deck_path, int argc,
                           char** argv, double sd = 1.0, double eps_l = 0.0) {
    {
        std::ifstream in(deck_path, std::ios::binary);
        std::ofstream out(outdir + "/deck_snapshot.ini", std::ios::binary);
        out << in.rdbuf();
    }
    char stamp[64] = {0};
    const std::time_t now = std::time(nullptr);
    std::strftime(stamp, sizeof stamp, "%Y-%m-%d %H:%M:%S %Z", std::localtime(&now));
    std::ofstream m(outdir + "/RUN_META.txt");
    m << "timestamp = " << stamp << "\n";
    m << "git_hash  = " << git_hash_best_effort() << "\n";
    m << "deck      = " << deck_path << " (snapshot: deck_snapshot.ini)\n";
    m << "S_d       = " << sd << "\n";
    m << "eps_L     = " << eps_l << "\n";
    m << "cmdline   =";
    for (int i = 0; i < argc; ++i) m << ' ' << argv[i];
    m << "\n";
}

}  // namespace arc

#endif  // ARC_PIC_RUN_META_HPP
