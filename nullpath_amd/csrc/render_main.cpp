// render_main.cpp -- the `render` command line: a drop-in for the reference
// renderer's executable (render.cu:391-647).  Kept from the reference, as its
// interface: every flag's spelling, default and value type; unknown flags and
// flags missing their value are ignored; --mode, --tonemap and --bg are
// case-folded; -h is the height; the validation messages, their order and
// exit code 1; the "[checkpoint] ..." and "Wrote ..." lines; the PPM bytes
// (truncating 8-bit tone curve of render.cu:583-607); --frames/--turns/--outpat
// animation with its one-%d pattern rule.
//
// The program is built around libnullpath.so's progressive passes: the
// accumulation buffer stays in device memory for the whole frame, the passes
// between two checkpoints are one replay of a captured HIP graph
// (np_pass_graph_*), and only checkpoints copy the image back.
//
// Additive flags (defaults reproduce the reference):
//   --precision fp64|fp32c|fp64-carter    --rtol X --atol X    --max-steps N    --seed N
//   --disk-rin isco|X         --disk-rout X        (X in units of M)
//   --emissivity-index q      (emissivity (r/M)^-q; default r^-2 absolute)
//   --raw-out FILE            (linear float32 RGB at the end, accum / spp, row-major)
//   --state FILE              (at every checkpoint: FILE = float32 accumulation sums,
//                              FILE.json = size, samples done, parameter signature)
//   --resume FILE             (continue a single-frame render from a --state file)
//   --fp64-radius X           (fp32c: FP64 steps inside X M; default max(30, 1.5 disk-rout))
//   --waves-per-simd N        (kernel variant: waves per SIMD, 0 = default)
//   --crossing linear|hermite (disk-crossing radius; default linear as the reference)
//   --estimate-every-step     (fp32c: no far-field skip of the error estimate)
//   --renorm-pr               (fp32c: p_r from H = 0 after accepted steps)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "nullpath.h"

namespace {

constexpr float kPi = 3.14159265358979323846f;

// ---------------------------------------------------------------------------
// Settings, grouped by what they control.
// ---------------------------------------------------------------------------
struct Settings {
    struct {
        int width = 1280, height = 720, tile_rows = 128;
    } image;
    struct {
        float fov_deg = 50.0f, theta_deg = 90.0f, phi0_deg = 0.0f, r_cam_rs = 50.0f;
        float mass = 10.0f, spin = 0.0f;
    } camera;
    struct {
        int per_pass = 1, total = 0, checkpoint_every = 0;
    } samples;
    struct {
        int frames = 1;
        float turns = 1.0f;
        std::string pattern = "frame_%05d.ppm";
        std::string single = "render.ppm";
    } output;
    struct {
        std::string mode = "balanced", tonemap = "none", background = "gradient";
        float exposure_ev = 0.0f, gamma = 2.2f, bg_sigma_deg = 1.0f, bg_brightness = 3.0f;
        float disk_gain = 1.0f;
        bool disk = true, invariants = false;
        int disk_orders = 1;
    } look;
    struct {   // additive extensions
        std::string precision = "fp64", crossing = "linear", disk_rin, raw_out, state, resume;
        double rtol = -1.0, atol = -1.0, disk_rout_m = -1.0, fp64_radius_m = 0.0;
        float emissivity = -1.0f;
        unsigned seed = 0;
        int max_steps = 10000, waves_per_simd = 0;
        bool estimate_every_step = false;
        bool renorm_pr = false;
    } ext;
};

// ---------------------------------------------------------------------------
// Flag table: each flag names the setting it writes; one loop applies them.
// ---------------------------------------------------------------------------
using Apply = std::function<void(Settings&, const char*)>;

struct Flag {
    std::vector<std::string> names;
    bool takes_value;
    Apply apply;
};

int as_int(const char* v) { return (int)std::strtol(v, nullptr, 10); }
float as_float(const char* v) { return std::strtof(v, nullptr); }
double as_double(const char* v) { return std::strtod(v, nullptr); }

std::vector<Flag> flag_table() {
    using S = Settings;
    std::vector<Flag> t;
    auto val = [&t](std::vector<std::string> names, Apply a) { t.push_back({std::move(names), true, std::move(a)}); };
    auto sw = [&t](std::vector<std::string> names, Apply a) { t.push_back({std::move(names), false, std::move(a)}); };
    val({"--w", "-w"}, [](S& s, const char* v) { s.image.width = as_int(v); });
    val({"--h", "-h"}, [](S& s, const char* v) { s.image.height = as_int(v); });
    val({"--tile"}, [](S& s, const char* v) { s.image.tile_rows = as_int(v); });
    val({"--fov"}, [](S& s, const char* v) { s.camera.fov_deg = as_float(v); });
    val({"--theta"}, [](S& s, const char* v) { s.camera.theta_deg = as_float(v); });
    val({"--phi0"}, [](S& s, const char* v) { s.camera.phi0_deg = as_float(v); });
    val({"--rcam"}, [](S& s, const char* v) { s.camera.r_cam_rs = as_float(v); });
    val({"--mass"}, [](S& s, const char* v) { s.camera.mass = as_float(v); });
    val({"--spin"}, [](S& s, const char* v) { s.camera.spin = as_float(v); });
    val({"--samples", "-s"}, [](S& s, const char* v) { s.samples.per_pass = as_int(v); });
    val({"--spp-total"}, [](S& s, const char* v) { s.samples.total = as_int(v); });
    val({"--checkpoint-every"}, [](S& s, const char* v) { s.samples.checkpoint_every = as_int(v); });
    val({"--frames"}, [](S& s, const char* v) { s.output.frames = as_int(v); });
    val({"--turns"}, [](S& s, const char* v) { s.output.turns = as_float(v); });
    val({"--outpat"}, [](S& s, const char* v) { s.output.pattern = v; });
    val({"--out"}, [](S& s, const char* v) { s.output.single = v; });
    val({"--mode"}, [](S& s, const char* v) { s.look.mode = v; });
    val({"--tonemap"}, [](S& s, const char* v) { s.look.tonemap = v; });
    val({"--bg"}, [](S& s, const char* v) { s.look.background = v; });
    val({"--exposure"}, [](S& s, const char* v) { s.look.exposure_ev = as_float(v); });
    val({"--gamma"}, [](S& s, const char* v) { s.look.gamma = as_float(v); });
    val({"--bg-sigma"}, [](S& s, const char* v) { s.look.bg_sigma_deg = as_float(v); });
    val({"--bg-bright"}, [](S& s, const char* v) { s.look.bg_brightness = as_float(v); });
    val({"--disk-gain"}, [](S& s, const char* v) { s.look.disk_gain = as_float(v); });
    val({"--disk-max-orders"}, [](S& s, const char* v) { s.look.disk_orders = as_int(v); });
    sw({"--no-disk"}, [](S& s, const char*) { s.look.disk = false; });
    sw({"--check-invariants"}, [](S& s, const char*) { s.look.invariants = true; });
    // extensions
    val({"--precision"}, [](S& s, const char* v) { s.ext.precision = v; });
    val({"--crossing"}, [](S& s, const char* v) { s.ext.crossing = v; });
    val({"--rtol"}, [](S& s, const char* v) { s.ext.rtol = as_double(v); });
    val({"--atol"}, [](S& s, const char* v) { s.ext.atol = as_double(v); });
    val({"--disk-rin"}, [](S& s, const char* v) { s.ext.disk_rin = v; });
    val({"--disk-rout"}, [](S& s, const char* v) { s.ext.disk_rout_m = as_double(v); });
    val({"--emissivity-index"}, [](S& s, const char* v) { s.ext.emissivity = as_float(v); });
    val({"--raw-out"}, [](S& s, const char* v) { s.ext.raw_out = v; });
    val({"--state"}, [](S& s, const char* v) { s.ext.state = v; });
    val({"--resume"}, [](S& s, const char* v) { s.ext.resume = v; });
    val({"--seed"}, [](S& s, const char* v) { s.ext.seed = (unsigned)std::strtoul(v, nullptr, 10); });
    val({"--max-steps"}, [](S& s, const char* v) { s.ext.max_steps = as_int(v); });
    val({"--fp64-radius"}, [](S& s, const char* v) { s.ext.fp64_radius_m = as_double(v); });
    val({"--waves-per-simd"}, [](S& s, const char* v) { s.ext.waves_per_simd = as_int(v); });
    sw({"--estimate-every-step"}, [](S& s, const char*) { s.ext.estimate_every_step = true; });
    sw({"--renorm-pr"}, [](S& s, const char*) { s.ext.renorm_pr = true; });
    return t;
}

Settings parse(int argc, char** argv) {
    const std::vector<Flag> table = flag_table();
    std::unordered_map<std::string, const Flag*> by_name;
    for (const Flag& f : table)
        for (const std::string& n : f.names) by_name[n] = &f;
    Settings s;
    for (int i = 1; i < argc; ++i) {
        const auto hit = by_name.find(argv[i]);
        if (hit == by_name.end()) continue;           // unknown: ignored
        const Flag& f = *hit->second;
        if (!f.takes_value) {
            f.apply(s, nullptr);
        } else if (i + 1 < argc) {
            f.apply(s, argv[++i]);                     // the next word, whatever it looks like
        }                                              // value missing: ignored
    }
    return s;
}

std::string folded(std::string v) {
    std::transform(v.begin(), v.end(), v.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return v;
}

// Validation rules in the order the reference applies them; the first one
// that fails ends the program with its message and status 1.
struct Rule {
    std::function<bool(const Settings&)> ok;
    const char* message;
};

const std::vector<Rule>& rules() {
    static const std::vector<Rule> r = {
        {[](const Settings& s) { return s.image.width > 0 && s.image.height > 0; }, "Invalid image dimensions."},
        {[](const Settings& s) { return s.image.tile_rows > 0; }, "Invalid tile height."},
        {[](const Settings& s) { return s.samples.per_pass > 0; }, "Invalid samples per pixel."},
        {[](const Settings& s) { return s.samples.total >= 0; }, "Invalid total spp."},
        {[](const Settings& s) { return s.output.frames > 0; }, "Invalid frame count."},
        {[](const Settings& s) { return s.camera.mass > 0.0f && s.camera.r_cam_rs > 0.0f; },
         "Invalid mass or camera radius."},
        {[](const Settings& s) { return std::fabs(s.camera.spin) < 1.0f; }, "Spin must satisfy |a/M| < 1."},
        {[](const Settings& s) { return s.camera.fov_deg > 0.0f && s.camera.fov_deg < 179.0f; },
         "Invalid FOV; expected (0, 179) degrees."},
        {[](const Settings& s) {
             const std::string p = folded(s.ext.precision);
             return p == "fp64" || p == "fp32c" || p == "fp64-carter";
         },
         "Invalid --precision; expected fp64, fp32c or fp64-carter."},
        {[](const Settings& s) {
             const std::string c = folded(s.ext.crossing);
             return c == "linear" || c == "hermite";
         },
         "Invalid --crossing; expected linear or hermite."},
        {[](const Settings& s) { return s.ext.resume.empty() || s.output.frames == 1; },
         "--resume applies to single-frame renders."},
    };
    return r;
}

// An --outpat must contain exactly one integer conversion: '%', an optional
// '-', optional digits, 'd'; "%%" is a literal percent sign.
bool one_int_conversion(const std::string& pat) {
    int conversions = 0;
    size_t i = 0;
    while (i < pat.size()) {
        if (pat[i] != '%') {
            ++i;
            continue;
        }
        if (i + 1 == pat.size()) return false;
        if (pat[i + 1] == '%') {
            i += 2;
            continue;
        }
        size_t j = i + 1;
        if (pat[j] == '-') ++j;
        while (j < pat.size() && std::isdigit((unsigned char)pat[j])) ++j;
        if (j == pat.size() || pat[j] != 'd') return false;
        ++conversions;
        i = j + 1;
    }
    return conversions == 1;
}

// Exposure, optional Reinhard curve, gamma, clamp, truncation to 8 bits --
// the float operations of render.cu:583-607 in the same order.
struct ToneCurve {
    float exposure_scale = 1.0f, inv_gamma = 1.0f;
    bool reinhard = false, gamma_on = false;

    ToneCurve(const Settings& s) {
        exposure_scale = std::pow(2.0f, s.look.exposure_ev);
        reinhard = folded(s.look.tonemap) == "reinhard";
        const float g = s.look.gamma <= 0.0f ? 1.0f : s.look.gamma;
        gamma_on = std::fabs(g - 1.0f) > 1e-6f;
        inv_gamma = 1.0f / std::fmax(1e-6f, g);
    }
    unsigned char operator()(float v) const {
        v *= exposure_scale;
        if (reinhard) v = v / (1.0f + v);
        if (gamma_on) v = std::pow(std::fmax(v, 0.0f), inv_gamma);
        v = std::fmin(std::fmax(v, 0.0f), 1.0f);
        return (unsigned char)(255.0f * v);
    }
};

bool save_ppm(const std::string& path, int w, int h, const std::vector<unsigned char>& pixels) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream head;
    head << "P6\n" << w << ' ' << h << "\n255\n";
    const std::string hs = head.str();
    f.write(hs.data(), (std::streamsize)hs.size());
    f.write(reinterpret_cast<const char*>(pixels.data()), (std::streamsize)pixels.size());
    return (bool)f;
}

bool save_floats(const std::string& path, const float* v, size_t n) {
    std::ofstream f(path, std::ios::binary);
    if (!f) return false;
    f.write(reinterpret_cast<const char*>(v), (std::streamsize)(n * sizeof(float)));
    return (bool)f;
}

// FNV-1a over the bytes of the render parameters up to the last field (the
// tail padding is not part of the value): a resumed render must trace
// exactly the rays the saved one did.
uint64_t signature(const np_render_params& p) {
    uint64_t h = 1469598103934665603ull;
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&p);
    const size_t n = offsetof(np_render_params, renorm_pr) + sizeof(p.renorm_pr);
    for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

// ---------------------------------------------------------------------------
// One frame's progressive render, accumulated on the device.
// ---------------------------------------------------------------------------
class ProgressiveFrame {
public:
    ProgressiveFrame(const np_render_params& p, const Settings& s) : p_(p), s_(s), curve_(s) {
        n_ = (size_t)p.width * (size_t)p.height * 3;
    }
    ~ProgressiveFrame() {
        if (graph_) np_pass_graph_destroy(graph_);
        if (d_accum_) (void)hipFree(d_accum_);
        if (d_inv_) (void)hipFree(d_inv_);
    }

    // Renders into `path`; false after printing the reason.
    bool run(const std::string& path) {
        const int per = std::max(1, s_.samples.per_pass);
        const int total = s_.samples.total > 0 ? s_.samples.total : s_.samples.per_pass;
        const int passes = (total + per - 1) / per;
        const int ckpt = s_.samples.checkpoint_every;
        if (hipMalloc(&d_accum_, n_ * sizeof(float)) != hipSuccess ||
            hipMalloc(&d_inv_, 3 * sizeof(float)) != hipSuccess ||
            hipMemset(d_inv_, 0, 3 * sizeof(float)) != hipSuccess)
            return fail("Render failed: device memory");
        int pass = 0;   // passes done
        if (!s_.ext.resume.empty()) {
            if (!load_state(s_.ext.resume, per, &pass)) return false;
            if (pass >= passes) return write_checkpoint(path, total, true);
        } else if (hipMemset(d_accum_, 0, n_ * sizeof(float)) != hipSuccess) {
            return fail("Render failed: device memory");
        }
        // full passes are replayed from a captured graph, group by group up
        // to the next checkpoint; a final short pass runs on its own
        const int full = total / per;
        const int group = ckpt > 0 ? ckpt : std::max(1, full);
        while (pass < passes) {
            const int next_ckpt = (ckpt > 0) ? std::min(passes, (pass / ckpt + 1) * ckpt) : passes;
            int err = NP_OK;
            if (pass + group <= std::min(full, next_ckpt) && (pass % group) == 0) {
                if (!graph_) err = np_pass_graph_create(&p_, per, group, d_accum_, inv_ptr(), pass * per, &graph_);
                if (err == NP_OK) err = np_pass_graph_seek(graph_, pass * per, nullptr);
                if (err == NP_OK) err = np_pass_graph_launch(graph_, nullptr);
                pass += group;
            } else {
                const int spp = std::min(per, total - pass * per);
                err = np_render_pass(&p_, pass * per, spp, d_accum_, inv_ptr(), nullptr);
                pass += 1;
            }
            if (err == NP_OK && hipDeviceSynchronize() != hipSuccess) err = NP_ELAUNCH;
            if (err != NP_OK) return fail(std::string("Render failed: ") + np_status_string(err));
            const bool checkpoint = pass == passes || (ckpt > 0 && pass % ckpt == 0);
            if (checkpoint && !write_checkpoint(path, std::min(total, pass * per), pass == passes)) return false;
        }
        return true;
    }

private:
    float* inv_ptr() { return p_.check_inv ? d_inv_ : nullptr; }

    bool fail(const std::string& what) {
        std::fprintf(stderr, "%s\n", what.c_str());
        return false;
    }

    bool write_checkpoint(const std::string& path, int spp_done, bool last) {
        std::vector<float> accum(n_);
        float inv[3] = {0.0f, 0.0f, 0.0f};
        if (hipMemcpy(accum.data(), d_accum_, n_ * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
            (p_.check_inv && hipMemcpy(inv, d_inv_, sizeof(inv), hipMemcpyDeviceToHost) != hipSuccess))
            return fail("Render failed: copy back");
        const float scale = 1.0f / (float)spp_done;
        std::vector<unsigned char> bytes(n_);
        for (size_t i = 0; i < n_; ++i) bytes[i] = curve_(accum[i] * scale);
        if (!save_ppm(path, p_.width, p_.height, bytes)) return fail("Failed to write " + path);
        if (p_.check_inv)
            std::printf("[checkpoint] %s @ %d spp  H|max=%.3e  dE/E|max=%.3e  dL/L|max=%.3e\n", path.c_str(),
                        spp_done, inv[0], inv[1], inv[2]);
        else
            std::printf("[checkpoint] %s @ %d spp\n", path.c_str(), spp_done);
        if (!s_.ext.state.empty() && !save_state(s_.ext.state, accum, spp_done)) return false;
        if (last && !s_.ext.raw_out.empty()) {
            std::vector<float> lin(n_);
            for (size_t i = 0; i < n_; ++i) lin[i] = accum[i] * scale;
            if (!save_floats(s_.ext.raw_out, lin.data(), n_)) return fail("Failed to write " + s_.ext.raw_out);
        }
        return true;
    }

    bool save_state(const std::string& path, const std::vector<float>& accum, int spp_done) {
        if (!save_floats(path, accum.data(), accum.size())) return fail("Failed to write " + path);
        std::ofstream j(path + ".json");
        j << "{\"width\": " << p_.width << ", \"height\": " << p_.height << ", \"spp_done\": " << spp_done
          << ", \"signature\": \"" << std::hex << signature(p_) << std::dec << "\"}\n";
        if (!j) return fail("Failed to write " + path + ".json");
        return true;
    }

    // Restores the sums and samples of a --state file; the render continues
    // at the next pass (sample ids continue, so the result equals an
    // uninterrupted render).
    bool load_state(const std::string& path, int per, int* pass) {
        std::ifstream j(path + ".json");
        std::stringstream text;
        text << j.rdbuf();
        const std::string js = text.str();
        auto field = [&js](const char* key) -> std::string {
            const std::string k = std::string("\"") + key + "\": ";
            const size_t at = js.find(k);
            if (at == std::string::npos) return "";
            size_t b = at + k.size(), e = b;
            if (b < js.size() && js[b] == '"') e = js.find('"', ++b);
            else
                while (e < js.size() && (std::isdigit((unsigned char)js[e]) || js[e] == '-')) ++e;
            return e == std::string::npos ? "" : js.substr(b, e - b);
        };
        std::ostringstream want;
        want << std::hex << signature(p_);
        const int done = std::atoi(field("spp_done").c_str());
        if (!j || std::atoi(field("width").c_str()) != p_.width ||
            std::atoi(field("height").c_str()) != p_.height || field("signature") != want.str())
            return fail("Cannot resume from " + path + ": it belongs to a different render");
        if (done <= 0 || done % per != 0)
            return fail("Cannot resume from " + path + ": its samples do not end on a pass of this render");
        std::vector<float> accum(n_);
        std::ifstream f(path, std::ios::binary);
        f.read(reinterpret_cast<char*>(accum.data()), (std::streamsize)(n_ * sizeof(float)));
        if (!f || f.gcount() != (std::streamsize)(n_ * sizeof(float)))
            return fail("Cannot resume from " + path + ": short file");
        if (hipMemcpy(d_accum_, accum.data(), n_ * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
            return fail("Render failed: device memory");
        *pass = done / per;
        return true;
    }

    np_render_params p_;
    const Settings& s_;
    ToneCurve curve_;
    size_t n_ = 0;
    float* d_accum_ = nullptr;
    float* d_inv_ = nullptr;
    np_pass_graph* graph_ = nullptr;
};

// The library's render description of these settings (render.cu:467-521
// conversions: degrees to radians in float, mode to integrator and camera).
np_render_params to_params(Settings& s) {
    auto rad = [](float d) { return d * kPi / 180.0f; };
    np_render_params p;
    np_render_params_default(&p);
    p.width = s.image.width;
    p.height = s.image.height;
    p.fov_y_deg = s.camera.fov_deg;
    p.theta_cam = rad(s.camera.theta_deg);
    p.mass = s.camera.mass;
    p.spin = s.camera.spin;
    p.r_cam_Rs = s.camera.r_cam_rs;
    if (folded(s.look.mode) == "accurate") {
        p.integrator = NP_INTEGRATOR_ADAPTIVE;
        p.camera = NP_CAMERA_ZAMO;
        s.look.invariants = true;
    }
    p.bg_mode = folded(s.look.background) == "point" ? NP_BG_POINT : NP_BG_GRADIENT;
    p.bg_sigma_rad = s.look.bg_sigma_deg * kPi / 180.0f;
    p.bg_peak = s.look.bg_brightness;
    p.disk_on = s.look.disk ? 1 : 0;
    p.disk_gain = s.look.disk_gain;
    p.disk_max_orders = std::max(1, s.look.disk_orders);
    p.check_inv = s.look.invariants ? 1 : 0;
    // extensions
    p.max_steps = s.ext.max_steps;
    const std::string prec = folded(s.ext.precision);
    p.precision = prec == "fp32c" ? NP_PREC_FP32C : prec == "fp64-carter" ? NP_PREC_FP64_CARTER : NP_PREC_FP64;
    if (s.ext.rtol > 0.0) p.rtol = s.ext.rtol;
    if (s.ext.atol >= 0.0) p.atol = s.ext.atol;
    p.seed = s.ext.seed;
    const double M = (double)s.camera.mass;
    p.fp64_radius = s.ext.fp64_radius_m > 0.0 ? s.ext.fp64_radius_m * M : 0.0;
    p.waves_per_simd = s.ext.waves_per_simd;
    p.crossing = folded(s.ext.crossing) == "hermite" ? NP_CROSSING_HERMITE : NP_CROSSING_LINEAR;
    p.estimate_every_step = s.ext.estimate_every_step ? 1 : 0;
    p.renorm_pr = s.ext.renorm_pr ? 1 : 0;
    if (!s.ext.disk_rin.empty())
        p.disk_rin = folded(s.ext.disk_rin) == "isco" ? np_isco_radius((double)s.camera.spin) * M
                                                       : std::strtod(s.ext.disk_rin.c_str(), nullptr) * M;
    if (s.ext.disk_rout_m > 0.0) p.disk_rout = s.ext.disk_rout_m * M;
    if (s.ext.emissivity >= 0.0f) {
        p.emiss_q = s.ext.emissivity;
        p.emiss_m_units = 1;
    }
    return p;
}

}  // namespace

int main(int argc, char** argv) {
    Settings s = parse(argc, argv);
    for (const Rule& r : rules())
        if (!r.ok(s)) {
            std::fprintf(stderr, "%s\n", r.message);
            return 1;
        }
    s.image.tile_rows = std::min(s.image.tile_rows, s.image.height);
    np_render_params p = to_params(s);

    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices <= 0) {
        std::fprintf(stderr, "No HIP devices found.\n");
        return 1;
    }
    (void)hipSetDevice(0);

    const float phi0 = s.camera.phi0_deg * kPi / 180.0f;
    if (s.output.frames <= 1) {
        p.phi_cam = phi0;
        ProgressiveFrame frame(p, s);
        if (!frame.run(s.output.single)) return 1;
        std::printf("Wrote %s\n", s.output.single.c_str());
        return 0;
    }
    if (!one_int_conversion(s.output.pattern)) {
        std::fprintf(stderr, "Invalid --outpat (expected a single %%d-like specifier).\n");
        return 1;
    }
    for (int i = 0; i < s.output.frames; ++i) {
        // azimuth of frame i, in the reference's float operation order
        const float t = (float)i / (float)s.output.frames;
        p.phi_cam = phi0 + 2.0f * kPi * s.output.turns * t;
        char path[512];
        std::snprintf(path, sizeof(path), s.output.pattern.c_str(), i);
        ProgressiveFrame frame(p, s);
        if (!frame.run(path)) return 1;
    }
    std::printf("Wrote %d frames matching pattern %s\n", s.output.frames, s.output.pattern.c_str());
    return 0;
}
