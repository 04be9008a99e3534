// driver.cpp — the `mlfma` executable: the reference's main() flow
// (src/mlfma.cpp:12-110) on top of the C-ABI.  Reads config/config.txt,
// config/pwave.txt and config/rwg/n<N>/{vertices,faces,rwgs}.txt (or a dipole
// file), builds the plan on the GPU, solves, and writes out/sol/curr_nq7.txt
// (one "re im" line per unknown at %.15e, src/solver.cpp:149-165).
//
// usage: mlfma [--root DIR] [--device D] [--method compat|gmres|fgmres|ir] [--restart M]
//              [--deflate K] [--maxit K] [--tol T] [--max-level L] [--dipoles FILE] [--fp32]
//        mlfma --root DIR --print-config   (print the parsed config.txt and exit;
//                                           no device is touched)
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "mlfma_c.h"

namespace fs = std::filesystem;

namespace {

// config.txt (src/config.h:16-58): one field per line, in the order of the
// Config constructor.  A line is read from its first digit on; a typed value
// is extracted, the stream skips to the next digit and extracts again, so the
// last number on the line wins (a sign is skipped like any other non-digit;
// an int field stops at a decimal point and takes the digits after it as the
// next number; a line without a number gives 0).  The extraction is the
// standard stream's operator>> for the field's type, as in the reference.
struct RefConfigFields {
    int mode = 0, pdist = 0, qdist = 0, quadPrec = 0, digits = 0, interpOrder = 0;
    double overInterp = 0, rootLeng = 0;
    int maxNodeSrcs = 0;
    bool evalDirect = false;
    int nsrcs = 0;
};

void skipToDigit(std::istringstream& iss) {
    char ch;
    while (iss.get(ch))
        if (std::isdigit(static_cast<unsigned char>(ch))) {
            iss.unget();
            break;
        }
}

template <typename T>
void readField(std::ifstream& f, T& val, int& lines) {
    std::string line;
    if (!std::getline(f, line)) return;
    ++lines;
    std::istringstream iss(line);
    skipToDigit(iss);
    T v{};  // a line without a number gives 0 (the reference leaves the field unset)
    while (iss >> v) skipToDigit(iss);
    val = v;
}

RefConfigFields readConfig(const fs::path& p) {
    std::ifstream f(p);
    if (!f) throw std::runtime_error("Unable to find file " + p.string());
    RefConfigFields c;
    int lines = 0;
    readField(f, c.mode, lines);
    readField(f, c.pdist, lines);
    readField(f, c.qdist, lines);
    readField(f, c.quadPrec, lines);
    readField(f, c.digits, lines);
    readField(f, c.interpOrder, lines);
    readField(f, c.overInterp, lines);
    readField(f, c.rootLeng, lines);
    readField(f, c.maxNodeSrcs, lines);
    readField(f, c.evalDirect, lines);
    readField(f, c.nsrcs, lines);
    if (lines < 11) throw std::runtime_error("config.txt needs 11 lines");
    return c;
}

std::vector<double> readRows(const fs::path& p, int cols) {
    std::ifstream f(p);
    if (!f) throw std::runtime_error("Unable to find file " + p.string());
    std::vector<double> v;
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream iss(line);
        double x;
        int c = 0;
        while (iss >> x) {
            v.push_back(x);
            ++c;
        }
        if (c == 0) continue;
        if (c < cols) throw std::runtime_error("Unable to parse line in " + p.string());
        for (; c > cols; --c) v.pop_back();
    }
    return v;
}

void check(int rc, const char* what) {
    if (rc != MLFMA_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, mlfma_last_error());
        std::exit(1);
    }
}

}  // namespace

int main(int argc, char** argv) {
    fs::path root = ".";
    int device = 0, restart = 50, maxit = 500, maxLevel = -1, deflate = 0;
    double tol = 1.0e-6;
    std::string method = "compat", dipoles;
    bool fp32 = false, printConfig = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) throw std::runtime_error("missing value for " + a);
            return argv[++i];
        };
        if (a == "--root") root = next();
        else if (a == "--device") device = std::stoi(next());
        else if (a == "--method") method = next();
        else if (a == "--restart") restart = std::stoi(next());
        else if (a == "--deflate") deflate = std::stoi(next());
        else if (a == "--maxit") maxit = std::stoi(next());
        else if (a == "--tol") tol = std::stod(next());
        else if (a == "--max-level") maxLevel = std::stoi(next());
        else if (a == "--dipoles") dipoles = next();
        else if (a == "--fp32") fp32 = true;
        else if (a == "--print-config") printConfig = true;
        else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 2;
        }
    }
    try {
        const RefConfigFields rc = readConfig(root / "config" / "config.txt");
        if (printConfig) {
            std::printf("mode %d\npdist %d\nqdist %d\nquadPrec %d\ndigits %d\ninterpOrder %d\n"
                        "overInterp %.17g\nrootLeng %.17g\nmaxNodeSrcs %d\nevalDirect %d\nnsrcs %d\n",
                        rc.mode, rc.pdist, rc.qdist, rc.quadPrec, rc.digits, rc.interpOrder, rc.overInterp,
                        rc.rootLeng, rc.maxNodeSrcs, rc.evalDirect ? 1 : 0, rc.nsrcs);
            return 0;
        }
        std::cout << " Importing sources...\n";
        mlfma_config cfg{};
        cfg.quad_prec = rc.quadPrec;
        cfg.digits = rc.digits;
        cfg.interp_order = rc.interpOrder;
        cfg.over_interp = rc.overInterp;
        cfg.root_leng = rc.rootLeng;
        cfg.max_node_srcs = rc.maxNodeSrcs;
        const bool evalDirect = rc.evalDirect;
        const int nsrcs = rc.nsrcs;
        cfg.max_level = maxLevel;
        cfg.device = device;
        cfg.precision = fp32 ? MLFMA_FP32 : MLFMA_FP64;

        const auto pwv = readRows(root / "config" / "pwave.txt", 8);
        mlfma_planewave pw{};
        pw.amplitude = pwv.at(0);
        for (int j = 0; j < 3; ++j) {
            pw.pol[j] = pwv.at(1 + j);
            pw.khat[j] = pwv.at(4 + j);
        }
        pw.wavenum = pwv.at(7);

        mlfma_geometry g{};
        std::vector<double> verts, pos, mom;
        std::vector<int32_t> tris, rwgs;
        if (!dipoles.empty()) {
            auto d = readRows(dipoles, 6);
            const int64_t n = (int64_t)d.size() / 6;
            for (int64_t i = 0; i < n; ++i)
                for (int j = 0; j < 3; ++j) {
                    pos.push_back(d[6 * i + j]);
                    mom.push_back(d[6 * i + 3 + j]);
                }
            g.kind = MLFMA_SRC_DIPOLE;
            g.n = n;
            g.dip_pos = pos.data();
            g.dip_mom = mom.data();
        } else {
            const fs::path dir = root / "config" / "rwg" / ("n" + std::to_string(nsrcs));
            verts = readRows(dir / "vertices.txt", 3);
            for (double v : readRows(dir / "faces.txt", 3)) tris.push_back((int32_t)v);
            for (double v : readRows(dir / "rwgs.txt", 4)) rwgs.push_back((int32_t)v);
            g.kind = MLFMA_SRC_RWG;
            g.n = (int64_t)rwgs.size() / 4;
            g.nverts = (int64_t)verts.size() / 3;
            g.verts = verts.data();
            g.ntris = (int64_t)tris.size() / 3;
            g.tris = tris.data();
            g.rwgs = rwgs.data();
        }
        std::cout << "   # Sources:       " << g.n << "\n   Digit precision: " << cfg.digits
                  << "\n   Interp order:    " << cfg.interp_order << "\n   Max node srcs:   " << cfg.max_node_srcs
                  << "\n   Root length:     " << cfg.root_leng << "\n   Wave number:     " << pw.wavenum << "\n\n";

        auto solveAndWrite = [&](mlfma_config c, const char* fname, const char* label) {
            std::cout << " Setting up " << label << " (tree, near field, tables, radiation patterns)...\n";
            auto t0 = std::chrono::steady_clock::now();
            mlfma_plan* plan = nullptr;
            check(mlfma_plan_create(&g, &pw, &c, &plan), "mlfma_plan_create");
            mlfma_info info{};
            check(mlfma_plan_info(plan, &info), "mlfma_plan_info");
            std::cout << "   # Nodes: " << info.num_nodes << "\n   # Leaves: " << info.num_leaves
                      << "\n   Max node level: " << info.max_level << "\n   (Lvl,Nth,Nph) =\n";
            for (int l = 0; l <= info.max_level; ++l)
                std::cout << "   (" << l << "," << info.nth[l] << "," << info.nph[l] << ")\n";
            std::cout << "   Elapsed time: "
                      << std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count()
                      << " ms\n\n Solving...\n";
            mlfma_solve_opts o{};
            o.method = method == "compat"   ? MLFMA_GMRES_COMPAT
                       : method == "fgmres" ? MLFMA_FGMRES_IR
                       : method == "ir"     ? MLFMA_GMRES_IR
                                            : MLFMA_GMRES_RESTART;
            o.restart = restart;
            o.deflate = deflate;  // gmres / ir: harmonic Ritz vectors kept per restart
            o.maxit = maxit;
            o.use_voltage = 1;
            o.tol = tol;
            mlfma_solve_report rep{};
            std::vector<double> x(2 * g.n), hist(maxit);
            check(mlfma_solve(plan, nullptr, x.data(), &o, &rep, hist.data()), "mlfma_solve");
            std::cout << "   Iterations: " << rep.iterations << "  residual estimate " << rep.resid_est
                      << "  true residual " << rep.resid_true << "\n   Total elapsed time: " << rep.seconds * 1e3
                      << " ms\n\n";
            fs::create_directories("out/sol");
            FILE* f = std::fopen((fs::path("out/sol") / fname).string().c_str(), "w");
            if (!f) throw std::runtime_error("cannot write out/sol");
            for (int64_t i = 0; i < g.n; ++i) std::fprintf(f, "%.15e %.15e\n", x[2 * i], x[2 * i + 1]);
            std::fclose(f);
            mlfma_plan_destroy(plan);
        };
        std::cout << " Solving w/ FMM...\n";
        solveAndWrite(cfg, "curr_nq7.txt", "FMM plan");
        if (evalDirect) {
            // the reference's direct path: a single root leaf (src/mlfma.cpp:84-107)
            mlfma_config d = cfg;
            d.max_node_srcs = (int32_t)g.n;
            d.max_level = 0;
            std::cout << " Solving w/ direct...\n";
            solveAndWrite(d, "currDir_nq7.txt", "direct plan");
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
