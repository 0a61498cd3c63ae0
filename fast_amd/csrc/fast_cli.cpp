// `fast tau` and `fast alpha` command line on top of libfasttau's C ABI.
//
// Mirrors the reference CLI (fast/src/main.cpp:20-134, fast/src/TauModule.h:12-164,
// fast/src/AlphaModule.h:8-219,
// fast/src/VolumeIOModule.h, fast/src/CSVOutputModule.h): the same flags, the same
// direction strings (fast/src/utility.h:84-113), one CSV row per direction with the same
// key-sorted columns and default 6-significant-digit formatting, header only when the
// output is new.  Inputs as loadVolume (fast/src/VolumeIOModule.h:41-131): a `.bin` file, a
// folder of TIFF slices (binarized with threshold 1), or a `.sph` sphere list rasterized at
// the --sub resolution; then the --sub/--origin, --randsub and --aspect crops
// (VolumeIOModule.h:153-245).  Additions: `-p mixed` (FP32 storage / FP64 arithmetic),
// `--reach FILE` (reachable-pore mask of the first direction's axis as .bin) and `--gpu N`.
// `--cpu` is not offered: this library has no CPU solver; `.pos` inputs are not read.
// `fast alpha` writes the reference's rows (AlphaModule::getAlpha, AlphaModule.h:28-174):
// per direction the boundary-connected area of that face, then a DIR_NONE row with the
// whole volume's area; with --basic, the whole volume's area on every direction's row.
// --segmentinfo FILE adds segmentVolumeRatio to the rows and writes the segments that do not
// reach every face to FILE (AlphaModule.h:36-43,65-123).  --mesh FILE writes the triangle mesh
// of the zero-padded volume as OBJ (AlphaModule.h:145-167).
#include "fast_tau.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <random>
#include <sys/stat.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

namespace {

const char* kDirNames[6] = {"X_POS", "X_NEG", "Y_POS", "Y_NEG", "Z_POS", "Z_NEG"};

std::vector<int> get_dirs(const std::string& s) {
    if (s == "all") return {0, 1, 2, 3, 4, 5};
    if (s == "pos") return {0, 2, 4};
    if (s == "neg") return {1, 3, 5};
    if (s == "xneg") return {1};
    if (s == "yneg") return {3};
    if (s == "zneg") return {5};
    std::vector<int> d;
    for (char c : s) {
        if (c == 'x') d.push_back(0);
        else if (c == 'y') d.push_back(2);
        else if (c == 'z') d.push_back(4);
    }
    if (d.empty()) return {0};
    return d;
}

// A CSV row: values kept as preformatted strings in a key-sorted map (std::map like the
// reference's RunRow), numbers formatted by a default ostream.
using Row = std::map<std::string, std::string>;
template <class V>
std::string fmt(const V& v) {
    std::ostringstream o;
    o << v;
    return o.str();
}

class CsvOut {
public:
    explicit CsvOut(const std::string& path) : path_(path) {}
    void add(const Row& row) {
        bool is_new;
        std::ostream* out = &std::cout;
        if (!path_.empty()) {
            std::ifstream probe(path_);
            is_new = !probe.is_open();
            if (!f_.is_open()) f_.open(path_, std::ios::app);
            out = &f_;
        } else {
            is_new = (rows_ == 0);
        }
        if (is_new) {
            size_t c = 0;
            for (auto& kv : row) *out << kv.first << (++c != row.size() ? "," : "");
            *out << "\n";
        }
        size_t c = 0;
        for (auto& kv : row) *out << kv.second << (++c != row.size() ? "," : "");
        *out << "\n";
        out->flush();
        ++rows_;
    }

private:
    std::string path_;
    std::ofstream f_;
    int rows_ = 0;
};

// "x" or "x,y,z" (the reference's ivec3/vec3 flag values)
template <class T>
bool parse3(const std::string& s, T (&v)[3]) {
    std::stringstream ss(s);
    std::string part;
    int n = 0;
    while (std::getline(ss, part, ',')) {
        if (n == 3) return false;
        std::istringstream ps(part);
        if (!(ps >> v[n])) return false;
        ++n;
    }
    if (n == 1) v[1] = v[2] = v[0];
    return n == 1 || n == 3;
}

bool is_directory(const std::string& p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

std::string extension(const std::string& p) {   // checkExtension: text after the last '.'
    const size_t dot = p.find_last_of('.');
    return dot == std::string::npos ? p : p.substr(dot + 1);
}

bool crop(std::vector<uint8_t>& m, int (&dim)[3], const int (&o)[3], const int (&sz)[3]) {
    std::vector<uint8_t> out((size_t)sz[0] * sz[1] * sz[2]);
    if (fast_crop(m.data(), dim[0], dim[1], dim[2], o[0], o[1], o[2], sz[0], sz[1], sz[2], out.data()) != FAST_OK)
        return false;
    m.swap(out);
    for (int a = 0; a < 3; ++a) dim[a] = sz[a];
    return true;
}

void usage() {
    std::cout << "Usage: fast tau|alpha <input> [options]   (input: .bin file, folder of .tif/.tiff slices, .sph)\n"
                 "  -d, --dir STR        x|y|z (combinable), all, pos, neg, xneg, yneg, zneg (default x-)\n"
                 "  --solver CG|BICGSTAB  (default CG)\n"
                 "  --tol N              tolerance 1e-N (default 6)\n"
                 "  --iter N             max iterations (default 10000)\n"
                 "  -p, --precision P    double|float|mixed (default double)\n"
                 "  -o, --output FILE    CSV output (append; header when new)\n"
                 "  -v, --verbose\n"
                 "  --over, --oversubscribe  accepted, no effect (288 GB HBM)\n"
                 "  --reach FILE         write the reachable-pore mask (.bin) of the first direction\n"
                 "  --gpu N              HIP device (default 0)\n"
                 "  --sub X|X,Y,Z        subvolume size (and the .sph resolution)\n"
                 "  --origin X|X,Y,Z     subvolume origin\n"
                 "  --randsub X|X,Y,Z    subvolume of this size at a random origin\n"
                 "  --aspect X|X,Y,Z     aspect-ratio crop; one component must be 1.0\n"
                 "alpha:\n"
                 "  --basic              whole-volume area only (no connectivity)\n"
                 "  --segmentinfo FILE   segments that do not reach every face, as CSV; adds\n"
                 "                       segmentVolumeRatio to the rows\n"
                 "  --mesh FILE          triangle mesh of the zero-padded volume as OBJ\n";
}

// AlphaModule::getAlpha (AlphaModule.h:28-174): porosity as getPorosity<float>, the
// per-face tensor (CCL included in its time, a sixth per row), then the DIR_NONE row.
// filterSegmentInfo (VolumeSegmentation.cpp:182-205) with the inverted mask: labels >= 1 not
// on row `row` of the per-face table (row 6: on every face).
std::vector<const fast_segment_info*> filter_segments(const std::vector<fast_segment_info>& sinfo,
                                                      const std::vector<uint8_t>& table, int row) {
    const size_t nl = sinfo.size();
    std::vector<const fast_segment_info*> out;
    for (size_t l = 1; l < nl; ++l)
        if (!table[(size_t)row * nl + l]) out.push_back(&sinfo[l]);
    return out;
}

// getCCLSegmentRatio (VolumeSegmentation.cpp:207-215)
float segment_ratio(const std::vector<const fast_segment_info*>& f, size_t n) {
    int vox = 0;
    for (auto* s : f) vox += (int)s->voxel_num;
    return float(vox) / (float)(int)n;
}

int run_alpha(fast_ctx* ctx, fast_volume* vol, const std::string& input, const int (&dim)[3], uint64_t zeros,
              bool basic, const std::vector<int>& dirs, const std::string& output, const std::string& seginfo) {
    CsvOut out(output);
    Row o;
    o["ID"] = input;
    o["dimx"] = fmt(dim[0]);
    o["dimy"] = fmt(dim[1]);
    o["dimz"] = fmt(dim[2]);
    const size_t n = (size_t)dim[0] * dim[1] * dim[2];
    o["porosity"] = fmt((float)zeros / (float)n);
    auto timed = [](auto&& f) {
        const auto t0 = std::chrono::system_clock::now();
        const int rc = f();
        return std::make_pair(rc, std::chrono::duration<double>(std::chrono::system_clock::now() - t0).count());
    };
    double whole = 0.0;
    if (!basic) {
        double tensor[6];
        auto [rc, dt] = timed([&] { return fast_alpha_area_tensor(vol, 0.5f, tensor); });
        if (rc != FAST_OK) {
            std::cerr << "alpha failed: " << fast_ctx_last_error(ctx) << std::endl;
            return 1;
        }
        std::vector<fast_segment_info> sinfo;
        std::vector<uint8_t> table;
        if (!seginfo.empty()) {
            uint32_t nl = 0;
            if (fast_ccl_labels(vol, 255, nullptr, &nl) != FAST_OK) {
                std::cerr << "segment info failed: " << fast_ctx_last_error(ctx) << std::endl;
                return 1;
            }
            sinfo.resize(nl);
            table.resize((size_t)7 * nl);
            if (fast_ccl_segment_info(vol, 255, sinfo.data(), nl) != FAST_OK ||
                fast_ccl_boundary_table(vol, 255, table.data(), nl) != FAST_OK) {
                std::cerr << "segment info failed: " << fast_ctx_last_error(ctx) << std::endl;
                return 1;
            }
        }
        for (int d : dirs) {
            o["rad"] = fmt(tensor[d]);
            o["radTime"] = fmt(dt / 6.0f);
            o["dir"] = kDirNames[d];
            if (!sinfo.empty()) o["segmentVolumeRatio"] = fmt(segment_ratio(filter_segments(sinfo, table, d), n));
            out.add(o);
        }
        auto [rc2, dt2] = timed([&] { return fast_alpha_area(vol, 0.5f, &whole); });
        if (rc2 != FAST_OK) {
            std::cerr << "alpha failed: " << fast_ctx_last_error(ctx) << std::endl;
            return 1;
        }
        o["rad"] = fmt((float)whole);
        o["radTime"] = fmt(dt2);
        o["dir"] = "DIR_NONE";
        if (!sinfo.empty()) {
            const auto none = filter_segments(sinfo, table, 6);
            o["segmentVolumeRatio"] = fmt(segment_ratio(none, n));
            CsvOut info(seginfo);
            for (auto* si : none) {
                Row r;
                r["ID"] = input;
                r["minX"] = fmt(si->min_bb[0]);
                r["minY"] = fmt(si->min_bb[1]);
                r["minZ"] = fmt(si->min_bb[2]);
                r["maxX"] = fmt(si->max_bb[0]);
                r["maxY"] = fmt(si->max_bb[1]);
                r["maxZ"] = fmt(si->max_bb[2]);
                r["labelID"] = fmt(si->label_id);
                r["voxelNum"] = fmt((int)si->voxel_num);
                r["atBoundary"] = fmt(si->at_boundary);
                info.add(r);
            }
        }
        out.add(o);
        return 0;
    }
    auto [rc, dt] = timed([&] { return fast_alpha_area(vol, 0.5f, &whole); });
    if (rc != FAST_OK) {
        std::cerr << "alpha failed: " << fast_ctx_last_error(ctx) << std::endl;
        return 1;
    }
    for (size_t k = 0; k < dirs.size(); ++k) {
        o["radBasic"] = fmt((float)whole);
        o["radBasicTime"] = fmt(dt);
        out.add(o);
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        usage();
        return 0;
    }
    const std::string action = argv[1];
    if (action == "-h" || action == "--help") {
        usage();
        return 0;
    }
    const bool alpha = (action == "alpha");
    if (action != "tau" && !alpha) {
        std::cerr << "Unknown action." << std::endl;
        usage();
        return 1;
    }
    std::string input, dirs = "x-", solver = "CG", prec = "double", output, reach, seginfo, mesh;
    int tol = 6, gpu = 0;
    long iters = 10000;
    bool verbose = false;
    int sub[3] = {0, 0, 0}, origin[3] = {0, 0, 0}, randsub[3] = {0, 0, 0};
    float aspect[3] = {1.0f, 1.0f, 1.0f};
    bool has_sub = false, has_randsub = false, has_aspect = false, basic = false;
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&](const char* what) -> std::string {
            if (i + 1 >= argc) {
                std::cerr << "missing value for " << what << std::endl;
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "-h" || a == "--help") {
            usage();
            return 0;
        } else if (a == "-d" || a == "--dir") {
            dirs = next("--dir");
        } else if (a == "--solver") {
            solver = next("--solver");
        } else if (a == "--tol") {
            tol = std::atoi(next("--tol").c_str());
        } else if (a == "--iter") {
            iters = std::atol(next("--iter").c_str());
        } else if (a == "-p" || a == "--precision") {
            prec = next("--precision");
        } else if (a == "-o" || a == "--output") {
            output = next("--output");
        } else if (a == "-v" || a == "--verbose") {
            verbose = true;
        } else if (a == "--over" || a == "--oversubscribe") {
        } else if (a == "-j" || a == "--cputhreads") {
            (void)next("--cputhreads");
        } else if (a == "--cpu") {
            std::cerr << "--cpu: this build has no CPU solver; the tau solve runs on the GPU" << std::endl;
            return 1;
        } else if (alpha && a == "--basic") {
            basic = true;
        } else if (alpha && a == "--segmentinfo") {
            seginfo = next("--segmentinfo");
        } else if (alpha && a == "--mesh") {
            mesh = next("--mesh");
        } else if (a == "--reach") {
            reach = next("--reach");
        } else if (a == "--gpu") {
            gpu = std::atoi(next("--gpu").c_str());
        } else if (a == "--sub" || a == "--origin" || a == "--randsub") {
            int* dst = a == "--sub" ? sub : (a == "--origin" ? origin : randsub);
            int v[3];
            if (!parse3(next(a.c_str()), v)) {
                std::cerr << "invalid value for " << a << std::endl;
                return 1;
            }
            std::copy(v, v + 3, dst);
            has_sub |= (a == "--sub");
            has_randsub |= (a == "--randsub");
        } else if (a == "--aspect") {
            if (!parse3(next("--aspect"), aspect)) {
                std::cerr << "invalid value for --aspect" << std::endl;
                return 1;
            }
            has_aspect = true;
        } else if (!a.empty() && a[0] != '-' && input.empty()) {
            input = a;
        } else {
            std::cerr << "unknown option " << a << std::endl;
            return 1;
        }
    }
    if (input.empty()) {
        std::cerr << "Invalid input" << std::endl;
        return 1;
    }
    // load (loadVolume, VolumeIOModule.h:41-131)
    int dim[3] = {0, 0, 0};
    std::vector<uint8_t> mask;
    if (is_directory(input)) {
        if (fast_tiff_folder_dims(input.c_str(), &dim[0], &dim[1], &dim[2]) != FAST_OK) {
            std::cerr << "Failed to load: " << input << std::endl;
            return 1;
        }
        mask.resize((size_t)dim[0] * dim[1] * dim[2]);
        if (fast_load_tiff_folder(input.c_str(), mask.data(), mask.size()) != FAST_OK) {
            std::cerr << "Failed to load: " << input << std::endl;
            return 1;
        }
        fast_binarize(mask.data(), mask.size(), 1.0f);
    } else if (extension(input) == "bin") {
        if (fast_load_bin_dims(input.c_str(), &dim[0], &dim[1], &dim[2]) != FAST_OK) {
            std::cerr << "Failed to load: " << input << std::endl;
            return 1;
        }
        mask.resize((size_t)dim[0] * dim[1] * dim[2]);
        if (fast_load_bin(input.c_str(), mask.data(), mask.size()) != FAST_OK) {
            std::cerr << "Failed to load: " << input << std::endl;
            return 1;
        }
    } else if (extension(input) == "sph") {
        // the resolution is the --sub value (VolumeIOModule.h:152-154)
        if (!has_sub || sub[0] <= 0 || sub[1] <= 0 || sub[2] <= 0) {
            std::cerr << "Invalid volume: a .sph input needs its resolution as --sub" << std::endl;
            return 1;
        }
        std::copy(sub, sub + 3, dim);
        mask.resize((size_t)dim[0] * dim[1] * dim[2]);
        if (fast_rasterize_sph(input.c_str(), dim[0], dim[1], dim[2], mask.data()) != FAST_OK) {
            std::cerr << "Failed to load: " << input << std::endl;
            return 1;
        }
    } else {
        std::cerr << "Invalid volume: " << input << std::endl;
        return 1;
    }
    // crops (VolumeIOModule.h:159-238)
    if (has_sub || origin[0] || origin[1] || origin[2]) {
        int lo[3], hi[3], sz[3];
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(std::max(origin[a], 0), dim[a]);
            hi[a] = std::min(std::max(origin[a] + sub[a], 0), dim[a]);
            sz[a] = hi[a] - lo[a];
            ok = ok && hi[a] > lo[a];
        }
        if (!ok || !crop(mask, dim, lo, sz)) {
            std::cerr << "Invalid resize options" << std::endl;
            return 1;
        }
    } else if (has_randsub) {
        std::mt19937 rng((unsigned)std::chrono::system_clock::now().time_since_epoch().count());
        int lo[3];
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
            const int max_origin = dim[a] - randsub[a];
            ok = ok && max_origin > 0 && randsub[a] > 0;
            lo[a] = ok ? (int)(rng() % (unsigned)max_origin) : 0;
        }
        if (!ok || !crop(mask, dim, lo, randsub)) {
            std::cerr << "Invalid resize options" << std::endl;
            return 1;
        }
    }
    if (has_aspect) {
        int k = -1;
        for (int a = 0; a < 3 && k < 0; ++a)
            if (aspect[a] == 1.0f) k = a;
        if (k < 0) {
            std::cerr << "Invalid aspect options, one dimension must equal to 1.0" << std::endl;
            return 1;
        }
        int sz[3];
        const int zero[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a) sz[a] = std::min(std::max((int)((float)dim[k] * aspect[a]), 0), dim[a]);
        if (!crop(mask, dim, zero, sz)) {
            std::cerr << "Invalid aspect options" << std::endl;
            return 1;
        }
    }
    const int nx = dim[0], ny = dim[1], nz = dim[2];
    fast_ctx* ctx = nullptr;
    int rc = fast_ctx_create(gpu, 0, &ctx);
    if (rc != FAST_OK) {
        std::cerr << "no usable GPU (" << rc << ")" << std::endl;
        return 1;
    }
    fast_volume* vol = nullptr;
    rc = fast_volume_create(ctx, mask.data(), nx, ny, nz, &vol);
    if (rc != FAST_OK) {
        std::cerr << "volume upload failed: " << fast_ctx_last_error(ctx) << std::endl;
        fast_ctx_destroy(ctx);
        return 1;
    }
    uint64_t zeros = 0;
    fast_volume_count_zero(vol, &zeros);
    if (alpha) {
        int status = run_alpha(ctx, vol, input, dim, zeros, basic, get_dirs(dirs), output, seginfo);
        if (status == 0 && !mesh.empty()) {   // AlphaModule.h:145-167
            size_t nv = 0;
            if (fast_alpha_save_obj(vol, 0.5f, mesh.c_str(), &nv) != FAST_OK) {
                std::cerr << "mesh export failed: " << fast_ctx_last_error(ctx) << std::endl;
                status = 1;
            } else if (nv > 0) {
                std::cout << "Saving " << mesh << ", Vertices:" << nv << std::endl;
                std::cout << "Done." << std::endl;
            } else {
                std::cerr << "Has no triangles" << std::endl;
            }
        }
        fast_volume_destroy(vol);
        fast_ctx_destroy(ctx);
        return status;
    }
    const float porosity = (float)((double)zeros / (double)mask.size());   // TauModule.h:62

    fast_tau_params p;
    fast_tau_params_default(&p);
    p.solver = (solver == "BICGSTAB") ? FAST_BICGSTAB : FAST_CG;
    const bool single = (prec == "float" || prec == "single" || prec == "f" || prec == "s");
    const bool mixed = (prec == "mixed" || prec == "m");
    p.precision = mixed ? FAST_MIXED : (single ? FAST_FP32 : FAST_FP64);
    p.tolerance = std::pow(10.0, -tol);
    p.max_iter = iters;
    p.porosity = porosity;
    p.verbose = verbose;

    CsvOut out(output);
    Row row;
    row["ID"] = input;
    row["precision"] = mixed ? "mixed" : (single ? "single" : "double");
    row["device"] = "GPU";
    row["dimx"] = fmt(nx);
    row["dimy"] = fmt(ny);
    row["dimz"] = fmt(nz);
    const std::vector<int> dl = get_dirs(dirs);
    int status = 0;
    for (int d : dl) {
        p.dir = d;
        row["dir"] = kDirNames[d];
        if (verbose) std::cout << "Tau Direction " << d << std::endl;
        fast_tau_result r;
        const auto t0 = std::chrono::system_clock::now();
        rc = fast_tau_solve_volume(vol, &p, &r, nullptr);
        const double dt = std::chrono::duration<double>(std::chrono::system_clock::now() - t0).count();
        if (rc != FAST_OK && rc != FAST_E_NOCONV) {
            std::cerr << "solve failed: " << fast_ctx_last_error(ctx) << std::endl;
            status = 1;
        }
        row["tau"] = fmt(r.tau);
        row["tauTime"] = fmt(dt);
        row["tauSolver"] = (p.solver == FAST_BICGSTAB) ? "BICGSTAB" : "CG";
        row["tauNU"] = "0";
        row["tauTolerance"] = fmt(p.tolerance);
        row["porosity"] = fmt((double)porosity);
        row["tauIter"] = fmt((int)r.iterations);
        if (verbose) std::cout << "Elapsed: " << dt << "s (" << dt / 60.0 << "m)" << std::endl;
        out.add(row);
    }
    if (!reach.empty()) {
        std::vector<uint8_t> rm(mask.size());
        uint32_t nl = 0;
        rc = fast_reachable_mask(vol, dl[0], 255, rm.data(), &nl);
        if (rc != FAST_OK || fast_save_bin(reach.c_str(), rm.data(), nx, ny, nz) != FAST_OK) {
            std::cerr << "reachable mask failed: " << fast_ctx_last_error(ctx) << std::endl;
            status = 1;
        }
    }
    fast_volume_destroy(vol);
    fast_ctx_destroy(ctx);
    return status;
}
