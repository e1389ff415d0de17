// render_main.cpp -- the `render` command line, a drop-in for the reference
// renderer's CLI (render.cu:391-647): same flags, defaults, parsing quirks
// (unknown flags and flags missing their value are ignored; --mode, --tonemap
// and --bg are case-folded; -h is the height), validation messages, exit
// codes, progressive passes with PPM checkpoints, and frame animation.
// Tracing runs through libnullpath.so (np_render_accumulate).
//
// Additive flags (defaults reproduce the reference):
//   --precision fp64|fp32c   --rtol X --atol X    --max-steps N   --seed N
//   --disk-rin isco|X        --disk-rout X        (X in units of M)
//   --emissivity-index q     (emissivity (r/M)^-q; default r^-2 absolute)
//   --raw-out file.f32       (linear float32 RGB, accum / spp, row-major)
//   --fp64-radius X          (fp32c: FP64 steps inside X M; default max(30, 1.5 disk-rout))
//   --waves-per-simd 2|3|4   (fp32c kernel variant: 3 = capped at 168 VGPRs, the default;
//                             4 = capped at 128 VGPRs; 2 = the compiler's register budget)
//   --crossing linear|hermite (disk-crossing radius; default linear as the reference)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "nullpath.h"

namespace {

struct Options {
    int width = 1280;
    int height = 720;
    float fov_y_deg = 50.0f;
    float theta_deg = 90.0f;
    float mass = 10.0f;
    float r_cam = 50.0f;
    const char* out_path = "render.ppm";
    int samples = 1;
    int tile_h = 128;
    int frames = 1;
    float turns = 1.0f;
    float phi0_deg = 0.0f;
    const char* out_pat = "frame_%05d.ppm";
    float spin = 0.0f;
    int spp_total = 0;
    int checkpoint_every = 0;
    int check_invariants = 0;
    const char* mode = "balanced";
    float exposure_ev = 0.0f;
    float gamma = 2.2f;
    const char* tonemap = "none";
    const char* bg = "gradient";
    float bg_sigma_deg = 1.0f;
    float bg_brightness = 3.0f;
    int disk_on = 1;
    float disk_gain = 1.0f;
    int disk_max_orders = 1;
    // extensions
    const char* precision = "fp64";
    double rtol = -1.0, atol = -1.0;
    const char* disk_rin = nullptr;
    double disk_rout_M = -1.0;
    float emiss_q = -1.0f;
    const char* raw_out = nullptr;
    unsigned seed = 0;
    int max_steps = 10000;
    double fp64_radius_M = 0.0;
    int waves_per_simd = 0;
    const char* crossing = "linear";
};

void parse_args(int argc, char** argv, Options& p) {
    for (int i = 1; i < argc; ++i) {
        std::string a(argv[i]);
        auto nextf = [&](float& ref) { if (i + 1 < argc) ref = strtof(argv[++i], nullptr); };
        auto nextd = [&](double& ref) { if (i + 1 < argc) ref = strtod(argv[++i], nullptr); };
        auto nexti = [&](int& ref) { if (i + 1 < argc) ref = atoi(argv[++i]); };
        auto nexts = [&](const char*& ref) { if (i + 1 < argc) ref = argv[++i]; };
        if (a == "--w" || a == "-w") nexti(p.width);
        else if (a == "--h" || a == "-h") nexti(p.height);
        else if (a == "--fov") nextf(p.fov_y_deg);
        else if (a == "--theta") nextf(p.theta_deg);
        else if (a == "--mass") nextf(p.mass);
        else if (a == "--rcam") nextf(p.r_cam);
        else if (a == "--out") nexts(p.out_path);
        else if (a == "--samples" || a == "-s") nexti(p.samples);
        else if (a == "--spp-total") nexti(p.spp_total);
        else if (a == "--checkpoint-every") nexti(p.checkpoint_every);
        else if (a == "--spin") nextf(p.spin);
        else if (a == "--check-invariants") p.check_invariants = 1;
        else if (a == "--tile") nexti(p.tile_h);
        else if (a == "--frames") nexti(p.frames);
        else if (a == "--turns") nextf(p.turns);
        else if (a == "--phi0") nextf(p.phi0_deg);
        else if (a == "--outpat") nexts(p.out_pat);
        else if (a == "--mode" && i + 1 < argc) p.mode = argv[++i];
        else if (a == "--exposure") nextf(p.exposure_ev);
        else if (a == "--gamma") nextf(p.gamma);
        else if (a == "--tonemap" && i + 1 < argc) p.tonemap = argv[++i];
        else if (a == "--bg" && i + 1 < argc) p.bg = argv[++i];
        else if (a == "--bg-sigma") nextf(p.bg_sigma_deg);
        else if (a == "--bg-bright") nextf(p.bg_brightness);
        else if (a == "--no-disk") p.disk_on = 0;
        else if (a == "--disk-gain") nextf(p.disk_gain);
        else if (a == "--disk-max-orders") nexti(p.disk_max_orders);
        // extensions
        else if (a == "--precision" && i + 1 < argc) p.precision = argv[++i];
        else if (a == "--rtol") nextd(p.rtol);
        else if (a == "--atol") nextd(p.atol);
        else if (a == "--disk-rin") nexts(p.disk_rin);
        else if (a == "--disk-rout") nextd(p.disk_rout_M);
        else if (a == "--emissivity-index") nextf(p.emiss_q);
        else if (a == "--raw-out") nexts(p.raw_out);
        else if (a == "--seed" && i + 1 < argc) p.seed = (unsigned)strtoul(argv[++i], nullptr, 10);
        else if (a == "--max-steps") nexti(p.max_steps);
        else if (a == "--fp64-radius") nextd(p.fp64_radius_M);
        else if (a == "--waves-per-simd") nexti(p.waves_per_simd);
        else if (a == "--crossing" && i + 1 < argc) p.crossing = argv[++i];
    }
}

bool validate_out_pattern(const char* pat) {
    if (!pat) return false;
    int count = 0;
    for (const char* q = pat; *q; ++q) {
        if (*q != '%') continue;
        ++q;
        if (*q == '\0') return false;
        if (*q == '%') continue;
        if (*q == '-') ++q;
        while (*q >= '0' && *q <= '9') ++q;
        if (*q != 'd') return false;
        count++;
    }
    return count == 1;
}

bool write_ppm(const char* path, int w, int h, const std::vector<unsigned char>& rgb) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    std::fprintf(f, "P6\n%d %d\n255\n", w, h);
    const size_t wrote = std::fwrite(rgb.data(), 1, rgb.size(), f);
    std::fclose(f);
    return wrote == rgb.size();
}

std::string lower(const char* s) {
    std::string o(s ? s : "");
    for (auto& c : o) c = (char)std::tolower((unsigned char)c);
    return o;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    parse_args(argc, argv, o);
    // validation, in the reference's order and words (render.cu:431-465)
    if (o.width <= 0 || o.height <= 0) { std::fprintf(stderr, "Invalid image dimensions.\n"); return 1; }
    if (o.tile_h <= 0) { std::fprintf(stderr, "Invalid tile height.\n"); return 1; }
    if (o.tile_h > o.height) o.tile_h = o.height;
    if (o.samples <= 0) { std::fprintf(stderr, "Invalid samples per pixel.\n"); return 1; }
    if (o.spp_total < 0) { std::fprintf(stderr, "Invalid total spp.\n"); return 1; }
    if (o.frames <= 0) { std::fprintf(stderr, "Invalid frame count.\n"); return 1; }
    if (o.mass <= 0.0f || o.r_cam <= 0.0f) { std::fprintf(stderr, "Invalid mass or camera radius.\n"); return 1; }
    if (std::fabs(o.spin) >= 1.0f) { std::fprintf(stderr, "Spin must satisfy |a/M| < 1.\n"); return 1; }
    if (o.fov_y_deg <= 0.0f || o.fov_y_deg >= 179.0f) {
        std::fprintf(stderr, "Invalid FOV; expected (0, 179) degrees.\n");
        return 1;
    }
    const std::string prec = lower(o.precision);
    if (prec != "fp64" && prec != "fp32c") {
        std::fprintf(stderr, "Invalid --precision; expected fp64 or fp32c.\n");
        return 1;
    }
    const std::string crossing = lower(o.crossing);
    if (crossing != "linear" && crossing != "hermite") {
        std::fprintf(stderr, "Invalid --crossing; expected linear or hermite.\n");
        return 1;
    }

    // mode mapping (render.cu:467-474)
    const std::string mode = lower(o.mode);
    int integrator = NP_INTEGRATOR_FIXED, camera = NP_CAMERA_STATIC;
    if (mode == "accurate") {
        integrator = NP_INTEGRATOR_ADAPTIVE;
        camera = NP_CAMERA_ZAMO;
        o.check_invariants = 1;
    }

    int device_count = 0;
    if (hipGetDeviceCount(&device_count) != hipSuccess || device_count <= 0) {
        std::fprintf(stderr, "No HIP devices found.\n");
        return 1;
    }
    (void)hipSetDevice(0);

    const float pi_f = 3.14159265358979323846f;
    auto deg2rad = [pi_f](float d) { return d * pi_f / 180.0f; };

    np_render_params p;
    np_render_params_default(&p);
    p.width = o.width;
    p.height = o.height;
    p.fov_y_deg = o.fov_y_deg;
    p.theta_cam = deg2rad(o.theta_deg);
    p.mass = o.mass;
    p.spin = o.spin;
    p.r_cam_Rs = o.r_cam;
    p.integrator = integrator;
    p.camera = camera;
    p.bg_mode = (lower(o.bg) == "point") ? NP_BG_POINT : NP_BG_GRADIENT;
    p.bg_sigma_rad = o.bg_sigma_deg * pi_f / 180.0f;
    p.bg_peak = o.bg_brightness;
    p.disk_on = o.disk_on ? 1 : 0;
    p.disk_gain = o.disk_gain;
    p.disk_max_orders = std::max(1, o.disk_max_orders);
    p.check_inv = o.check_invariants;
    p.max_steps = o.max_steps;
    p.precision = (prec == "fp32c") ? NP_PREC_FP32C : NP_PREC_FP64;
    if (o.rtol > 0.0) p.rtol = o.rtol;
    if (o.atol >= 0.0) p.atol = o.atol;
    p.seed = o.seed;
    const double M = (double)o.mass;
    p.fp64_radius = o.fp64_radius_M > 0.0 ? o.fp64_radius_M * M : 0.0;
    p.waves_per_simd = o.waves_per_simd;
    p.crossing = (crossing == "hermite") ? NP_CROSSING_HERMITE : NP_CROSSING_LINEAR;
    p.disk_rin = 3.0 * 2.0 * M;
    p.disk_rout = 20.0 * 2.0 * M;
    if (o.disk_rin) {
        p.disk_rin = (lower(o.disk_rin) == "isco") ? np_isco_radius((double)o.spin) * M
                                                  : strtod(o.disk_rin, nullptr) * M;
    }
    if (o.disk_rout_M > 0.0) p.disk_rout = o.disk_rout_M * M;
    if (o.emiss_q >= 0.0f) {
        p.emiss_q = o.emiss_q;
        p.emiss_m_units = 1;
    }

    const int tonemap_mode = (lower(o.tonemap) == "reinhard") ? 1 : 0;
    const float exposure_scale = powf(2.0f, o.exposure_ev);
    const float gamma = (o.gamma <= 0.0f ? 1.0f : o.gamma);

    auto render_one = [&](float phi_cam, const char* path) -> bool {
        p.phi_cam = phi_cam;
        const int total_spp = (o.spp_total > 0 ? o.spp_total : o.samples);
        const int batch = std::max(1, o.samples);
        const int passes = (total_spp + batch - 1) / batch;
        std::vector<float> accum((size_t)o.width * o.height * 3, 0.0f);
        int spp_done = 0;
        float inv_max[3] = {0, 0, 0};
        for (int k = 0; k < passes; ++k) {
            const int pass_spp = std::min(batch, total_spp - spp_done);
            const int err = np_render_accumulate(&p, spp_done, pass_spp, 1, o.tile_h, accum.data(),
                                                 o.check_invariants ? inv_max : nullptr);
            if (err != NP_OK) {
                std::fprintf(stderr, "Render failed: %s\n", np_status_string(err));
                return false;
            }
            spp_done += pass_spp;
            const bool do_ckpt = (k == passes - 1) ||
                                 (o.checkpoint_every > 0 && ((k + 1) % o.checkpoint_every == 0));
            if (!do_ckpt) continue;
            // tone mapping of render.cu:583-607 (truncating 8-bit conversion)
            std::vector<unsigned char> out((size_t)o.width * o.height * 3);
            const float inv = 1.0f / (float)spp_done;
            for (size_t i = 0, n = out.size() / 3; i < n; ++i) {
                float c[3];
                for (int ch = 0; ch < 3; ++ch) {
                    float v = accum[3 * i + ch] * inv;
                    v *= exposure_scale;
                    if (tonemap_mode == 1) v = v / (1.0f + v);
                    if (std::fabs(gamma - 1.0f) > 1e-6f) v = powf(std::fmax(v, 0.0f), 1.0f / std::fmax(1e-6f, gamma));
                    c[ch] = std::fmin(std::fmax(v, 0.0f), 1.0f);
                }
                out[3 * i + 0] = (unsigned char)(255.0f * c[0]);
                out[3 * i + 1] = (unsigned char)(255.0f * c[1]);
                out[3 * i + 2] = (unsigned char)(255.0f * c[2]);
            }
            if (!write_ppm(path, o.width, o.height, out)) {
                std::fprintf(stderr, "Failed to write %s\n", path);
                return false;
            }
            if (o.check_invariants)
                std::printf("[checkpoint] %s @ %d spp  H|max=%.3e  dE/E|max=%.3e  dL/L|max=%.3e\n", path,
                            spp_done, inv_max[0], inv_max[1], inv_max[2]);
            else
                std::printf("[checkpoint] %s @ %d spp\n", path, spp_done);
            if (o.raw_out && k == passes - 1) {
                FILE* f = std::fopen(o.raw_out, "wb");
                bool ok = f != nullptr;
                for (size_t i = 0; ok && i < accum.size(); ++i) {
                    const float v = accum[i] * inv;
                    ok = std::fwrite(&v, sizeof(float), 1, f) == 1;
                }
                if (f) std::fclose(f);
                if (!ok) {
                    std::fprintf(stderr, "Failed to write %s\n", o.raw_out);
                    return false;
                }
            }
        }
        return true;
    };

    const float phi0 = deg2rad(o.phi0_deg);
    if (o.frames <= 1) {
        if (!render_one(phi0, o.out_path)) return 1;
        std::printf("Wrote %s\n", o.out_path);
    } else {
        if (!validate_out_pattern(o.out_pat)) {
            std::fprintf(stderr, "Invalid --outpat (expected a single %%d-like specifier).\n");
            return 1;
        }
        for (int i = 0; i < o.frames; ++i) {
            const float t = (float)i / (float)o.frames;
            const float phi = phi0 + 2.0f * pi_f * o.turns * t;
            char path[512];
            std::snprintf(path, sizeof(path), o.out_pat, i);
            if (!render_one(phi, path)) return 1;
        }
        std::printf("Wrote %d frames matching pattern %s\n", o.frames, o.out_pat);
    }
    return 0;
}
