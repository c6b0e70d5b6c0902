// litho_gpu: the reference's `litho <config.json> <output_dir>` flow (Modules
// 1-7 of validation/golden/src/litho_thread.cpp main, L1228-1531) on the
// MI355X engine, writing the same BIN set the reference writes:
//   mskf_r.bin / mskf_i.bin (L1391-1392), tcc_r.bin / tcc_i.bin (L1417-1418),
//   aerial_image_tcc_direct.bin (L1440), kernels/krn_{k}_{r,i}.bin,
//   kernels/kernel_info.txt, scales.bin (L1465-1473),
//   tmpImgp_full_{G}.bin (L1015), aerial_image_socs_kernel.bin (L1494),
//   fft_meta.txt (L1200-1226).
// and, like source/host/socs_host.cpp --load-kernels (L101-179), it can resume
// from an existing kernels/ directory instead of Modules 4-6.
//
// Plain C++ over the C-ABI (include/litho_socs.h): no HIP types here.
//
// Config: the reference's JSON schema (input/config/golden_1024.json) plus
// two optional fields: mask.pixelNm (nm per pixel, default 1 as in the
// reference, where the period in nm equals the pixel count) and
// mask.type "Manhattan" with mask.seed (seeded synthetic rectangles).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "litho_socs.h"

namespace fs = std::filesystem;

// ---------------------------------------------------------------- JSON ------
struct JVal {
    enum Kind { NUL, NUM, STR, BOOL, OBJ, ARR } kind = NUL;
    double num = 0;
    bool b = false;
    std::string str;
    std::map<std::string, JVal> obj;
    std::vector<JVal> arr;
};

class JParser {
   public:
    explicit JParser(const std::string& t) : s(t) {}
    JVal parse() {
        JVal v = value();
        ws();
        if (i != s.size()) fail("trailing characters");
        return v;
    }

   private:
    const std::string& s;
    size_t i = 0;
    [[noreturn]] void fail(const char* m) {
        throw std::runtime_error(std::string("config JSON: ") + m + " at offset " +
                                 std::to_string(i));
    }
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) i++;
    }
    JVal value() {
        ws();
        if (i >= s.size()) fail("unexpected end");
        char c = s[i];
        JVal v;
        if (c == '{') {
            v.kind = JVal::OBJ;
            i++;
            ws();
            if (s[i] == '}') {
                i++;
                return v;
            }
            for (;;) {
                ws();
                std::string k = string();
                ws();
                if (s[i++] != ':') fail("expected ':'");
                v.obj[k] = value();
                ws();
                if (s[i] == ',') {
                    i++;
                    continue;
                }
                if (s[i++] != '}') fail("expected '}'");
                return v;
            }
        }
        if (c == '[') {
            v.kind = JVal::ARR;
            i++;
            ws();
            if (s[i] == ']') {
                i++;
                return v;
            }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (s[i] == ',') {
                    i++;
                    continue;
                }
                if (s[i++] != ']') fail("expected ']'");
                return v;
            }
        }
        if (c == '"') {
            v.kind = JVal::STR;
            v.str = string();
            return v;
        }
        if (s.compare(i, 4, "true") == 0) {
            i += 4;
            v.kind = JVal::BOOL;
            v.b = true;
            return v;
        }
        if (s.compare(i, 5, "false") == 0) {
            i += 5;
            v.kind = JVal::BOOL;
            return v;
        }
        if (s.compare(i, 4, "null") == 0) {
            i += 4;
            return v;
        }
        char* end = nullptr;
        v.num = std::strtod(s.c_str() + i, &end);
        if (end == s.c_str() + i) fail("bad value");
        i = end - s.c_str();
        v.kind = JVal::NUM;
        return v;
    }
    std::string string() {
        if (s[i] != '"') fail("expected string");
        i++;
        std::string out;
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\' && i + 1 < s.size()) i++;
            out += s[i++];
        }
        if (i >= s.size()) fail("unterminated string");
        i++;
        return out;
    }
};

static const JVal* jget(const JVal& root, const std::string& path) {
    const JVal* v = &root;
    std::stringstream ss(path);
    std::string part;
    while (std::getline(ss, part, '.')) {
        if (v->kind != JVal::OBJ) return nullptr;
        auto it = v->obj.find(part);
        if (it == v->obj.end()) return nullptr;
        v = &it->second;
    }
    return v;
}
static double jnum(const JVal& r, const std::string& p, double d) {
    const JVal* v = jget(r, p);
    if (!v) return d;
    if (v->kind == JVal::BOOL) return v->b ? 1.0 : 0.0;
    return v->kind == JVal::NUM ? v->num : d;
}
static std::string jstr(const JVal& r, const std::string& p, const std::string& d) {
    const JVal* v = jget(r, p);
    return (v && v->kind == JVal::STR) ? v->str : d;
}

// ---------------------------------------------------------------- BIN I/O ---
static bool write_bin(const fs::path& p, const float* d, size_t n) {
    std::ofstream f(p, std::ios::binary);
    if (!f) return false;
    f.write(reinterpret_cast<const char*>(d), n * sizeof(float));
    return (bool)f;
}
static bool read_bin(const fs::path& p, std::vector<float>& d, size_t n) {
    std::ifstream f(p, std::ios::binary | std::ios::ate);
    if (!f) return false;
    size_t bytes = (size_t)f.tellg();
    if (bytes != n * sizeof(float)) {
        std::cerr << "error: " << p << " holds " << bytes << " bytes, expected "
                  << n * sizeof(float) << "\n";
        return false;
    }
    f.seekg(0);
    d.resize(n);
    f.read(reinterpret_cast<char*>(d.data()), bytes);
    return (bool)f;
}

// ---------------------------------------------------------------- helpers ---
struct Timer {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    long us() const {
        return (long)std::chrono::duration_cast<std::chrono::microseconds>(
                   std::chrono::steady_clock::now() - t0)
            .count();
    }
};

#define CHECK(call)                                                                   \
    do {                                                                              \
        int st_ = (call);                                                             \
        if (st_ != LITHO_OK) {                                                        \
            std::cerr << "error: " #call " -> " << litho_status_string(st_) << "\n"; \
            return 2;                                                                 \
        }                                                                             \
    } while (0)

static void usage(const char* a0) {
    std::cerr << "usage: " << a0
              << " <config.json> <output_dir> [--load-kernels <dir>] [--no-tcc] [--dry-run]"
                 " [--quiet|--verbose]\n"
                 "  --load-kernels  read kernels/ and scales from <dir> (skip Modules 4-6)\n"
                 "  --no-tcc        skip the explicit TCC and the TCC-direct image; kernels\n"
                 "                  come from the source-space route\n"
                 "  --dry-run       parse the config, print the derived parameters, exit\n";
}

int main(int argc, char** argv) {
    if (argc < 3) {
        usage(argv[0]);
        return 1;
    }
    std::string cfgPath = argv[1];
    fs::path out = argv[2];
    std::string loadDir;
    bool noTcc = false, dry = false;
    int verbose = 1;
    for (int a = 3; a < argc; a++) {
        std::string s = argv[a];
        if (s == "--load-kernels" && a + 1 < argc) loadDir = argv[++a];
        else if (s == "--no-tcc") noTcc = true;
        else if (s == "--dry-run") dry = true;
        else if (s == "--quiet") verbose = 0;
        else if (s == "--verbose" || s == "--debug") verbose = 2;
        else {
            usage(argv[0]);
            return 1;
        }
    }
    JVal cfg;
    try {
        std::ifstream f(cfgPath);
        if (!f) {
            std::cerr << "error: cannot open config " << cfgPath << "\n";
            return 1;
        }
        std::stringstream b;
        b << f.rdbuf();
        cfg = JParser(b.str()).parse();
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    // Module 0: parameters (loadConfig, L1099-1156)
    const int Lx = (int)jnum(cfg, "mask.period.Lx", 0), Ly = (int)jnum(cfg, "mask.period.Ly", 0);
    const double pix = jnum(cfg, "mask.pixelNm", 1.0);
    const int msx = (int)jnum(cfg, "mask.size.maskSizeX", 0);
    const int msy = (int)jnum(cfg, "mask.size.maskSizeY", 0);
    const std::string mtype = jstr(cfg, "mask.type", "LineSpace");
    const float dose = (float)jnum(cfg, "mask.dose", 1.0);
    const int srcSize = (int)jnum(cfg, "source.gridSize", 0);
    const std::string stype = jstr(cfg, "source.type", "Annular");
    const float NA = (float)jnum(cfg, "optics.NA", 0.0);
    const float lambda = (float)jnum(cfg, "optics.wavelength", 193.0);
    const float defocus = (float)jnum(cfg, "optics.defocus", 0.0);
    int nk = (int)jnum(cfg, "kernel.count", 10);
    if (Lx <= 0 || Ly <= 0 || msx <= 0 || msy <= 0 || srcSize <= 0 || NA <= 0.f || pix <= 0) {
        std::cerr << "error: invalid or unsupported mask/source/optics parameters\n";
        return 1;
    }
    // pixels per side (physLx x physLy, L1228-1233); Lx != Ly runs the
    // rectangular entries, which take power-of-two sides
    const int N = (int)std::lround(Lx / pix), NyP = (int)std::lround(Ly / pix);
    if (std::fabs(N * pix - Lx) > 1e-6 || std::fabs(NyP * pix - Ly) > 1e-6 || msx > N ||
        msy > NyP) {
        std::cerr << "error: period " << Lx << " x " << Ly << " nm is not a whole number of "
                  << pix << " nm pixels covering the mask\n";
        return 1;
    }
    const bool rect = NyP != N;
    // Module 1: source
    std::vector<float> src((size_t)srcSize * srcSize, 0.f);
    if (stype == "Annular")
        CHECK(litho_source_annular(src.data(), srcSize, (float)jnum(cfg, "source.annular.innerRadius", 0.6),
                                   (float)jnum(cfg, "source.annular.outerRadius", 0.9)));
    else if (stype == "Dipole")
        CHECK(litho_source_dipole(src.data(), srcSize, (float)jnum(cfg, "source.dipole.radius", 0.5),
                                  (float)jnum(cfg, "source.dipole.offset", 0.3),
                                  (int)jnum(cfg, "source.dipole.onXAxis", 1)));
    else if (stype == "CrossQuadrupole")
        CHECK(litho_source_quadrupole(src.data(), srcSize,
                                      (float)jnum(cfg, "source.crossQuadrupole.radius", 0.5),
                                      (float)jnum(cfg, "source.crossQuadrupole.offset", 0.3)));
    else if (stype == "Point")
        CHECK(litho_source_point(src.data(), srcSize, (float)jnum(cfg, "source.point.x", 0.0),
                                 (float)jnum(cfg, "source.point.y", 0.0)));
    else if (stype == "Import") {
        if (!read_bin(jstr(cfg, "source.inputFile", ""), src, src.size())) return 1;
        float sum = 0.f;
        for (float v : src) sum += v;
        if (sum != 0.f)
            for (float& v : src) v /= sum;
    } else {
        std::cerr << "error: unknown source type " << stype << "\n";
        return 1;
    }
    const float sigma = litho_outer_sigma(src.data(), srcSize);
    const int Nx = litho_compute_nx(NA, Lx, sigma, lambda);
    const int Ny = rect ? litho_compute_nx(NA, Ly, sigma, lambda) : Nx;  // L1300-1301
    const int W = 2 * Nx + 1, H = 2 * Ny + 1, n = W * H;
    int G = 1, Gy = 1;
    while (G < 4 * Nx + 1) G *= 2;
    while (Gy < 4 * Ny + 1) Gy *= 2;
    if (nk > n) nk = n;  // L1456
    if (verbose) {
        std::cout << "[CONFIG] period " << Lx;
        if (rect) std::cout << " x " << Ly;
        std::cout << " nm, " << N << "x" << NyP << " pixels of " << pix << " nm, NA " << NA
                  << ", lambda " << lambda << ", defocus " << defocus << "\n[PARAM] outer sigma "
                  << sigma;
        if (rect)
            std::cout << ", Nx = " << Nx << ", Ny = " << Ny;
        else
            std::cout << ", Nx = Ny = " << Nx;
        std::cout << ", TCC " << n << "x" << n << ", conv grid " << G;
        if (rect) std::cout << "x" << Gy;
        std::cout << ", kernels " << nk << "\n";
    }
    if (dry) return 0;
    fs::create_directories(out / "kernels");
    {
        std::ofstream m(out / "fft_meta.txt");
        m << "physical_size_x " << N << "\nphysical_size_y " << NyP << "\nfft_size_x " << N
          << "\nfft_size_y " << NyP << "\nNx " << Nx << "\nNy " << Ny << "\nconv_size_x "
          << 4 * Nx + 1 << "\nconv_size_y " << 4 * Ny + 1 << "\nfft_conv_size_x " << G
          << "\nfft_conv_size_y " << Gy << "\n";
    }
    // Module 2: mask (createMask, common/mask.cpp L31-60)
    std::vector<float> inner((size_t)msx * msy, 0.f), mask((size_t)N * NyP, 0.f);
    if (mtype == "LineSpace") {
        const int lw = (int)jnum(cfg, "mask.lineSpace.lineWidth", 128);
        const int sw = (int)jnum(cfg, "mask.lineSpace.spaceWidth", 128);
        const bool hor = jnum(cfg, "mask.lineSpace.isHorizontal", 0) != 0;
        for (int y = 0; y < msy; y++)
            for (int x = 0; x < msx; x++)
                inner[(size_t)y * msx + x] = ((hor ? y : x) % (lw + sw)) < lw ? 1.f : 0.f;
    } else if (mtype == "Import") {
        if (!read_bin(jstr(cfg, "mask.inputFile", ""), inner, inner.size())) return 1;
    } else if (mtype == "Manhattan") {
        if (msx != msy) {
            std::cerr << "error: Manhattan masks are square\n";
            return 1;
        }
        CHECK(litho_manhattan_mask(inner.data(), msx, (unsigned long long)jnum(cfg, "mask.seed", 1),
                                   (int)jnum(cfg, "mask.minWidth", 12),
                                   (int)jnum(cfg, "mask.maxWidth", 45),
                                   (float)jnum(cfg, "mask.fill", 0.4)));
    } else {
        std::cerr << "error: unknown mask type " << mtype << "\n";
        return 1;
    }
    const int oy = (NyP - msy) / 2, ox = (N - msx) / 2;
    for (int y = 0; y < msy; y++)
        for (int x = 0; x < msx; x++)
            mask[(size_t)(y + oy) * N + x + ox] = inner[(size_t)y * msx + x] * dose;

    long tm[8] = {0};
    // Module 3: mask spectrum
    std::vector<float> mr((size_t)N * NyP), mi((size_t)N * NyP);
    {
        Timer t;
        if (rect)
            CHECK(litho_mask_spectrum_rect(mask.data(), N, NyP, mr.data(), mi.data()));
        else
            CHECK(litho_mask_spectrum(mask.data(), N, mr.data(), mi.data()));
        tm[3] = t.us();
    }
    write_bin(out / "mskf_r.bin", mr.data(), mr.size());
    write_bin(out / "mskf_i.bin", mi.data(), mi.size());

    std::vector<float> kr((size_t)nk * n), ki((size_t)nk * n), sc(nk);
    std::vector<double> scd(nk);
    LithoOptics op{lambda, NA, defocus, sigma, Lx, Ly, Nx, Ny};
    if (loadDir.empty()) {
        std::vector<double> tr, ti;
        if (!noTcc) {
            // Module 4 + 5: TCC and the TCC-direct image
            tr.resize((size_t)n * n);
            ti.resize((size_t)n * n);
            Timer t;
            CHECK(litho_tcc_build(&op, src.data(), srcSize, tr.data(), ti.data()));
            tm[4] = t.us();
            std::vector<float> f32((size_t)n * n);
            for (size_t i = 0; i < f32.size(); i++) f32[i] = (float)tr[i];
            write_bin(out / "tcc_r.bin", f32.data(), f32.size());
            for (size_t i = 0; i < f32.size(); i++) f32[i] = (float)ti[i];
            write_bin(out / "tcc_i.bin", f32.data(), f32.size());
            const auto p2 = [](int v) { return v >= 256 && v <= 1024 && (v & (v - 1)) == 0; };
            if (p2(N) && p2(NyP)) {
                Timer t5;
                std::vector<float> img((size_t)N * NyP);
                if (rect)
                    CHECK(litho_image_tcc_rect(mask.data(), N, NyP, tr.data(), ti.data(), Nx, Ny,
                                               img.data()));
                else
                    CHECK(litho_image_tcc(mask.data(), N, tr.data(), ti.data(), Nx, Nx,
                                          img.data()));
                tm[5] = t5.us();
                write_bin(out / "aerial_image_tcc_direct.bin", img.data(), img.size());
            }
        }
        // Module 6: kernels
        int flags = 0;
        Timer t;
        if (!noTcc)
            CHECK(litho_kernels_topk(tr.data(), ti.data(), n, Nx, Ny, nk, scd.data(), kr.data(),
                                     ki.data(), &flags));
        else
            CHECK(litho_kernels_from_source(&op, src.data(), srcSize, nk, scd.data(), kr.data(),
                                            ki.data(), &flags));
        tm[6] = t.us();
        if (flags & LITHO_FLAG_CLUSTER_SPLIT)
            std::cout << "[WARN] kernel count " << nk
                      << " cuts a degenerate eigenvalue cluster; the kernels inside it are in "
                         "the canonical mirror basis\n";
        for (int k = 0; k < nk; k++) sc[k] = (float)scd[k];  // L645
        for (int k = 0; k < nk; k++) {
            write_bin(out / "kernels" / ("krn_" + std::to_string(k) + "_r.bin"), &kr[(size_t)k * n], n);
            write_bin(out / "kernels" / ("krn_" + std::to_string(k) + "_i.bin"), &ki[(size_t)k * n], n);
        }
        write_bin(out / "scales.bin", sc.data(), nk);
        std::ofstream info(out / "kernels" / "kernel_info.txt");
        info << "# Kernel Information\n## Kernel Size and Count\n- Kernel Size: " << W << "x" << H
             << "\n- Number of Kernels: " << nk << "\n## Eigenvalues of Each Kernel\n";
        for (int k = 0; k < nk; k++) info << "Eigenvalue " << k << ": " << sc[k] << "\n";
    } else {
        // resume (socs_host --load-kernels): scales.bin + kernels/krn_*
        fs::path d = loadDir;
        std::vector<float> s;
        if (!read_bin(d / "scales.bin", s, nk)) return 1;
        sc = s;
        for (int k = 0; k < nk; k++) {
            std::vector<float> r, i;
            if (!read_bin(d / "kernels" / ("krn_" + std::to_string(k) + "_r.bin"), r, n) ||
                !read_bin(d / "kernels" / ("krn_" + std::to_string(k) + "_i.bin"), i, n))
                return 1;
            std::copy(r.begin(), r.end(), kr.begin() + (size_t)k * n);
            std::copy(i.begin(), i.end(), ki.begin() + (size_t)k * n);
        }
    }
    // Module 7: SOCS (HLS-shaped tmpImgp and the full image)
    {
        Timer t;
        std::vector<float> tmp((size_t)G * Gy);
        CHECK(litho_socs_tmpimg_on_grid(mr.data(), mi.data(), kr.data(), ki.data(), sc.data(),
                                        tmp.data(), nullptr, nk, Nx, Ny, N, NyP, G, Gy));
        write_bin(out / ("tmpImgp_full_" + std::to_string(G) +
                         (rect ? "x" + std::to_string(Gy) : std::string()) + ".bin"),
                  tmp.data(), tmp.size());
        std::vector<float> img((size_t)N * NyP);
        if (rect)
            CHECK(litho_socs_image_rect(mask.data(), N, NyP, kr.data(), ki.data(), sc.data(), nk,
                                        Nx, Ny, img.data()));
        else
            CHECK(litho_socs_image(mask.data(), N, kr.data(), ki.data(), sc.data(), nk, Nx, Nx,
                                   img.data()));
        tm[7] = t.us();
        write_bin(out / "aerial_image_socs_kernel.bin", img.data(), img.size());
        if (verbose) {
            double mean = 0, mx = -1e30, mn = 1e30;
            for (float v : img) {
                mean += v;
                mx = std::max(mx, (double)v);
                mn = std::min(mn, (double)v);
            }
            std::cout << "[STATISTICS] Aerial Image (SOCS): mean " << mean / img.size() << " max "
                      << mx << " min " << mn << "\n";
        }
    }
    if (verbose) {
        std::cout << "[SUMMARY] Module 3 (FFT Mask) " << tm[3] << " us\n";
        if (tm[4]) std::cout << "[SUMMARY] Module 4 (TCC) " << tm[4] << " us\n";
        if (tm[5]) std::cout << "[SUMMARY] Module 5 (TCC Image) " << tm[5] << " us\n";
        if (tm[6]) std::cout << "[SUMMARY] Module 6 (Kernels) " << tm[6] << " us\n";
        std::cout << "[SUMMARY] Module 7 (SOCS Image) " << tm[7] << " us\n[DONE] Output saved to: "
                  << out.string() << "\n";
    }
    return 0;
}
