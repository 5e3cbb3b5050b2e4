// xenodon: the command-line front end (render / convert / help / sysinfo)
// on top of the MI355X library (include/xenodon.h).
//
// Follows src/main.cpp (subcommands, option set, messages), src/main_loop.cpp
// (frame loop, shader selection, log lines), src/render/RenderStats.cpp (the
// --stats-output text), src/backend/headless/ (config, PNG frames on a black
// canvas), src/convert.cpp and src/sysinfo.cpp.  Everything the reference
// reports through its Error type is printed as "Error: <message>" and the
// process exits with status 0, as the reference does.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "cli_util.hpp"
#include "volume_io.hpp"
#include "xenodon.h"

using namespace xdcli;

namespace {

void check(int status) {
    if (status != XD_OK) throw Error(xd_last_error());
}

// ------------------------------------------------------------------ help ---

const char* kHelp =
    "Usage:\n"
    "    xenodon [subcommand] [subcommand options]\n"
    "\n"
    "Subcommands:\n"
    "help [subcommand or topic]\n"
    "    Show help for a subcommand or topic; without arguments, this message.\n"
    "\n"
    "sysinfo\n"
    "    List the GPUs this build can render on.\n"
    "\n"
    "convert [options] <source> <destination>\n"
    "    Convert a 3D TIFF stack to a sparse voxel octree (.svo).\n"
    "\n"
    "render [options] <volume>\n"
    "    Render a volume.\n"
    "\n"
    "headless-config\n"
    "    The configuration format of the headless rendering backend.\n";

const char* kHelpHelp =
    "Usage:\n"
    "    xenodon help [subcommand or topic]\n"
    "\n"
    "Show the help message of Xenodon, or of one subcommand or topic.\n";

const char* kHelpSysinfo =
    "Usage:\n"
    "    xenodon sysinfo\n"
    "\n"
    "List the GPUs that xenodon can render on:\n"
    "\n"
    "GPU <gpu-index>:\n"
    "    name: <gpu-name>\n"
    "    type: <type>\n"
    "    pci bus: <domain>:<bus>:<device>.<function>\n"
    "\n"
    "<gpu-index> is the vkindex of a headless configuration's device entry.\n";

const char* kHelpConvert =
    "Usage:\n"
    "    xenodon convert [options] <source tiff path> <destination svo path>\n"
    "\n"
    "Convert the 3D TIFF stack at <source tiff path> (one page per z layer) into\n"
    "a sparse voxel octree at <destination svo path>.\n"
    "\n"
    "Options:\n"
    "--dag\n"
    "    Share equal subtrees: the tree becomes a directed acyclic graph.\n"
    "\n"
    "--rope\n"
    "    Give every leaf links (ropes) to its six face neighbours; required by\n"
    "    the svo-rope traversal, and accepted by every other SVO traversal.\n"
    "\n"
    "--chan-diff <value>\n"
    "    Merge a region into one leaf when, in every colour channel, its voxels\n"
    "    differ by at most <value> (0-255). --chan-diff 0 is lossless.\n"
    "\n"
    "--std-dev <value>\n"
    "    Merge a region into one leaf when the standard deviation of its voxel\n"
    "    colours around their average is at most <value>.\n"
    "\n"
    "--gpu <gpu-index>\n"
    "    Build on that GPU instead of the host (sparse and rope trees with the\n"
    "    channel-difference heuristic; the result is identical).\n";

const char* kHelpRender =
    "Usage:\n"
    "    xenodon render [options] <volume path>\n"
    "\n"
    "Render the volume at <volume path> (.tif/.tiff grid or .svo octree).\n"
    "\n"
    "Options:\n"
    "-q --quiet                 Do not log to standard output.\n"
    "--log-output <file>        Also log to <file>.\n"
    "--repeat <amount>          Render every camera frame <amount> times (default 1).\n"
    "--volume-type <type>       'tiff', 'tif' or 'svo' instead of the file extension.\n"
    "--camera <file>            Camera transforms, one per line:\n"
    "                           fx fy fz ux uy uz px py pz (forward, up, position);\n"
    "                           the volume spans (0,0,0) to (voxel ratio).\n"
    "-e --emission-coeff <e>    Emission coefficient (default 1).\n"
    "-s --shader <name>         dda (grids), svo-naive, svo-df, esvo or svo-rope\n"
    "                           (octrees; svo-rope needs a `convert --rope` tree).\n"
    "-r --voxel-ratio <x>:<y>:<z>  Scale of the volume (default 1:1:1).\n"
    "--stats-output <file>      Save per-frame statistics to <file>.\n"
    "\n"
    "Backend:\n"
    "--headless <config>        Render without a display; each frame is saved as a\n"
    "                           PNG image (see 'xenodon help headless-config').\n"
    "    --discard-output       Do not save frames.\n"
    "    --output <format>      Output path format, with the frame number as its\n"
    "                           argument; default 'out-{}.png' ('out-{:0>3}.png'\n"
    "                           pads to three digits).\n";

const char* kHelpHeadless =
    "The headless backend configuration file has this format:\n"
    "\n"
    "device {\n"
    "    vkindex = <gpu-index>\n"
    "    offset = (<offset x>, <offset y>)\n"
    "    extent = (<width>, <height>)\n"
    "}\n"
    "\n"
    "Any number of device entries may be given; each renders the region of the\n"
    "frame at <offset> of size <extent>, in pixels, on GPU <gpu-index> (see\n"
    "'xenodon sysinfo').  The saved frame is the union of all regions.\n";

void help(const char* prog, const std::vector<std::string>& args) {
    if (args.empty()) {
        fputs(kHelp, stdout);
        return;
    }
    if (args.size() != 1) {
        printf("Error: Invalid usage of subcommand 'help', see '%s help'\n", prog);
        return;
    }
    static const std::map<std::string, const char*> topics = {
        {"help", kHelpHelp}, {"sysinfo", kHelpSysinfo}, {"convert", kHelpConvert},
        {"render", kHelpRender}, {"headless-config", kHelpHeadless}};
    auto it = topics.find(args[0]);
    if (it == topics.end())
        printf("Error: No such topic %s, see '%s help'\n", args[0].c_str(), prog);
    else
        fputs(it->second, stdout);
}

// --------------------------------------------------------------- sysinfo ---

void sysinfo() {
    printf("System setup information\n");
    int n = 0;
    if (xd_device_count(&n) != XD_OK || n == 0) {
        printf("No GPUs detected\n");
        return;
    }
    for (int i = 0; i < n; ++i) {
        char name[256] = {};
        uint32_t dom = 0, bus = 0, dev = 0, fn = 0;
        uint64_t bytes = 0;
        if (xd_device_describe(i, name, sizeof name, &dom, &bus, &dev, &fn, &bytes) != XD_OK) continue;
        printf("GPU %d:\n\tname: '%s'\n\ttype: DiscreteGpu\n", i, name);
        // "{:0>8}:{:0>2}:{:0>2}.{}" of src/sysinfo.cpp:53-59: decimal fields, zero-padded
        printf("\tpci bus: %08u:%02u:%02u.%u\n", dom, bus, dev, fn);
        printf("\tNo displays detected\n");
    }
}

// --------------------------------------------------------------- convert ---

void convert(const std::vector<std::string>& args) {
    std::string src, dst;
    bool dag = false, rope = false;
    int64_t chan_diff = -1;
    double stddev = -1;
    int64_t gpu = -1;
    Command cmd;
    cmd.flags = {{&dag, "--dag"}, {&rope, "--rope"}};
    cmd.parameters = {
        {[&](const std::string& a) { int64_t v; if (!parse_int(a, v) || v < 0 || v > 1024) return false; gpu = v; return true; },
         "gpu index", "--gpu"},
        {[&](const std::string& a) { int64_t v; if (!parse_int(a, v) || v < 0 || v > 255) return false; chan_diff = v; return true; },
         "channel difference", "--chan-diff"},
        {[&](const std::string& a) { double v; if (!parse_float(a, v) || v < 0.0) return false; stddev = v; return true; },
         "std. dev", "--std-dev"}};
    cmd.positional = {{[&](const std::string& a) { src = a; return true; }, "source tiff path"},
                      {[&](const std::string& a) { dst = a; return true; }, "destination svo path"}};
    try {
        parse_args(args, cmd);
    } catch (const ParseError& e) {
        printf("Error: %s\n", e.what());
        return;
    }
    if (dag && rope) {
        printf("Error: --dag and --rope are mutually exclusive\n");
        return;
    }
    if (chan_diff >= 0 && stddev >= 0) {
        printf("Error: --std-dev and --chan-diff are mutually exclusive\n");
        return;
    }
    printf("Loading source...\n");
    fflush(stdout);
    GridData grid;
    try {
        grid = load_tiff(src);
    } catch (const Error& e) {
        printf("Error reading '%s': %s\n", src.c_str(), e.what());
        return;
    }
    printf("Source grid:\n");
    printf(" Dimensions: %llux%llux%llu\n", (unsigned long long)grid.dim[0], (unsigned long long)grid.dim[1],
           (unsigned long long)grid.dim[2]);
    printf(" Size: %s bytes\n", fmt_grouped(grid.texels.size() * 4).c_str());
    printf("Converting to octree...\n");
    fflush(stdout);
    uint32_t* nodes = nullptr;
    xd_build_stats st{};
    const int heuristic = stddev >= 0 ? XD_HEURISTIC_STD_DEV : XD_HEURISTIC_CHAN_DIFF;
    const int type = dag ? XD_TREE_DAG : rope ? XD_TREE_ROPE : XD_TREE_SPARSE;
    if (gpu >= 0) {
        // the grid goes to HBM once; the tree comes back as .svo words
        xd_context* ctx = nullptr;
        xd_volume* vol = nullptr;
        int rc = xd_open((int)gpu, &ctx);
        if (rc == XD_OK) rc = xd_volume_from_grid(ctx, grid.texels.data(), grid.dim[0], grid.dim[1], grid.dim[2], &vol);
        if (rc == XD_OK)
            rc = xd_build_octree_device(ctx, vol, heuristic, (int)std::max<int64_t>(chan_diff, 0), std::max(stddev, 0.0),
                                        type, &nodes, &st);
        const std::string msg = rc == XD_OK ? "" : xd_last_error();
        xd_close(ctx);
        if (rc != XD_OK) {
            printf("Error: %s\n", msg.c_str());
            return;
        }
    } else if (xd_build_octree(grid.texels.data(), grid.dim[0], grid.dim[1], grid.dim[2], heuristic,
                               (int)std::max<int64_t>(chan_diff, 0), std::max(stddev, 0.0), type, &nodes, &st) !=
               XD_OK) {
        printf("Error: %s\n", xd_last_error());
        return;
    }
    SvoData svo;
    svo.side = st.side;
    svo.words.assign(nodes, nodes + st.node_count * 10);
    xd_free(nodes);
    // convert.cpp:78-93 reports node counts against a "perfect tree" figure
    // computed as (8^depth - 1) / 9
    uint64_t pw = 1;
    for (uint64_t k = 1; k < st.depth + 1; ++k) pw *= 8;
    const uint64_t perfect = (pw - 1) / 9;
    const double pt = (double)perfect;
    printf("Generated octree:\n");
    printf(" Dimensions: %llux%llux%llu\n", (unsigned long long)st.side, (unsigned long long)st.side,
           (unsigned long long)st.side);
    printf(" Size: %s bytes\n", fmt_grouped(st.node_count * 40).c_str());
    printf(" Perfect tree nodes: %s\n", fmt_grouped(perfect).c_str());
    printf(" Total nodes: %s (%.5f%%)\n", fmt_grouped(st.total_nodes).c_str(), (double)st.total_nodes / pt * 100);
    printf(" Unique nodes: %s (%.5f%%)\n", fmt_grouped(st.node_count).c_str(), (double)st.node_count / pt * 100);
    printf(" Total leaves: %s\n", fmt_grouped(st.total_leaves).c_str());
    printf(" Unique leaves: %s\n", fmt_grouped(st.unique_leaves).c_str());
    printf(" Depth: %s\n", fmt_grouped(st.depth).c_str());
    try {
        save_svo(dst, svo);
    } catch (const Error& e) {
        printf("Error writing '%s': %s\n", dst.c_str(), e.what());
    }
}

// ---------------------------------------------------------------- render ---

struct RenderStats {  // src/render/RenderStats.h:13-27
    uint64_t total_rays = 0, outputs = 0;
    double total = 0, max = 0, min = std::numeric_limits<double>::max();
    double mrays() const { return (double)total_rays / (total * 1000.0); }
};

// RenderStatsAccumulator::save (src/render/RenderStats.cpp:127-158)
void save_stats(const std::string& path, const std::vector<RenderStats>& frames, double fps) {
    uint64_t rays = 0;
    double time = 0;
    for (const auto& f : frames) {
        rays += f.total_rays;
        time += f.total;
    }
    std::string out;
    out += sformat("total rays: %llu\n", (unsigned long long)rays);
    out += "total render time: " + fmt_double(time) + "\n";
    out += "total mray/s: " + fmt_double((double)rays / (time * 1000.0)) + "\n";
    out += "average fps: " + fmt_double(fps) + "\n";
    out += sformat("frames: %zu\n", frames.size());
    out += "# Frame number: total rays, outputs, total render time, max render time, min render time, mray/s\n";
    for (size_t i = 0; i < frames.size(); ++i) {
        const auto& f = frames[i];
        out += sformat("frame %zu: %llu rays, %llu, ", i, (unsigned long long)f.total_rays,
                       (unsigned long long)f.outputs) +
               fmt_double(f.total) + " ms, " + fmt_double(f.max) + " ms, " + fmt_double(f.min) + " ms, " +
               fmt_double(f.mrays()) + " mray/s\n";
    }
    std::ofstream o(path);
    if (!o) throw Error("Failed to open render stats output path '" + path + "'");
    o << out;
}

// ScriptCameraController (src/camera/ScriptCameraController.cpp:4-41): one
// transform per line; the loop ends after the frame of the last line.
class CameraScript {
    std::ifstream in_;

public:
    float v[9] = {};
    explicit CameraScript(const std::string& path) : in_(path) {
        if (!in_) throw Error("Failed to open '" + path + "'");
        update();
    }
    // true at end of input
    bool update() {
        if (in_.eof()) return true;
        for (float& x : v) in_ >> x;
        if (in_.fail()) throw Error("Syntax error in camera input file");
        in_ >> std::ws;
        return false;
    }
};

struct Output {
    HeadlessDevice cfg;
    xd_context* ctx = nullptr;
    xd_frame* frame = nullptr;
};

struct RenderOptions {
    bool quiet = false, xorg = false, discard = false;
    std::string log_output, headless, output, direct, xorg_multi_gpu, volume_type, shader, stats, camera, volume;
    double emission = 1.0;
    float ratio[3] = {1, 1, 1};
    int64_t repeat = 1;
};

RenderOptions parse_render_args(const std::vector<std::string>& args) {
    RenderOptions o;
    auto str = [](std::string* v) { return [v](const std::string& a) { *v = a; return true; }; };
    Command cmd;
    cmd.flags = {{&o.quiet, "--quiet", 'q'}, {&o.xorg, "--xorg"}, {&o.discard, "--discard-output"}};
    cmd.parameters = {
        {str(&o.log_output), "output path", "--log-output"},
        {str(&o.headless), "config path", "--headless"},
        {str(&o.output), "output path", "--output"},
        {str(&o.direct), "config path", "--direct"},
        {str(&o.xorg_multi_gpu), "config path", "--xorg-multi-gpu"},
        {[&](const std::string& a) { double v; if (!parse_float(a, v) || v < 0.0) return false; o.emission = v; return true; },
         "emission coefficient", "--emission-coeff", 'e'},
        {str(&o.volume_type), "volume type", "--volume-type"},
        {str(&o.shader), "shader", "--shader", 's'},
        {[&](const std::string& a) {
             const size_t f = a.find(':'), s = f == std::string::npos ? f : a.find(':', f + 1);
             if (f == std::string::npos || s == std::string::npos) return false;
             const std::string p[3] = {a.substr(0, f), a.substr(f + 1, s - f - 1), a.substr(s + 1)};
             double v[3];
             for (int i = 0; i < 3; ++i)
                 if (p[i].empty() || !parse_float(p[i], v[i])) return false;
             if (!(v[0] > 0 && v[1] > 0 && v[2] > 0)) return false;
             for (int i = 0; i < 3; ++i) o.ratio[i] = (float)v[i];
             return true;
         },
         "voxel dimension ratio", "--voxel-ratio", 'r'},
        {str(&o.stats), "stats output", "--stats-output"},
        {str(&o.camera), "camera", "--camera"},
        {[&](const std::string& a) { int64_t v; if (!parse_int(a, v) || v < 1 || v > 0x7FFFFFFF) return false; o.repeat = v; return true; },
         "frame repeat", "--repeat"}};
    cmd.positional = {{str(&o.volume), "volume path"}};
    parse_args(args, cmd);
    const int backends = (int)o.xorg + (int)!o.headless.empty() + (int)!o.direct.empty();
    if (backends == 0) throw Error("Missing required backend --xorg, --headless or --direct");
    if (backends > 1) throw Error("--xorg, --headless and --direct are mutually exclusive");
    if (o.headless.empty() && o.discard) throw Error("--dont-save requires --headless");
    if (!o.output.empty() && o.headless.empty()) throw Error("--output requires --headless");
    if (o.output.empty())
        o.output = "out-{}.png";
    else if (o.discard)
        throw Error("--dont-save and --output are mutually exclusive");
    if (!o.xorg_multi_gpu.empty() && !o.xorg) throw Error("--xorg-multi-gpu requires --xorg");
    return o;
}

std::string extension_type(const RenderOptions& o) {
    std::string t = o.volume_type;
    if (t.empty()) {
        const size_t slash = o.volume.find_last_of('/');
        const std::string base = slash == std::string::npos ? o.volume : o.volume.substr(slash + 1);
        const size_t dot = base.find_last_of('.');
        if (dot == std::string::npos || dot == 0) return "unknown";
        t = base.substr(dot + 1);
    }
    if (t == "tiff" || t == "tif") return "tiff";
    if (t == "svo") return "svo";
    return "unknown";
}

void main_loop(Logger& log, const RenderOptions& o, std::vector<Output>& outs) {
    // check_setup (main_loop.cpp:47-71): every headless device has one output
    const size_t n = outs.size();
    std::string setup = sformat("Setup: %zu %s, with ", n, n > 1 ? "devices" : "device");
    for (size_t i = 0; i < n; ++i) setup += i ? ", 1" : "1";
    setup += n > 1 ? " outputs" : " output";
    log.log(setup);

    const std::string type = extension_type(o);
    if (type == "unknown") throw Error("Failed to parse model file type");
    log.log("Model file type: '" + type + "'");
    static const char* kShaders[5] = {"dda", "svo-naive", "esvo", "svo-df", "svo-rope"};
    int traversal = type == "tiff" ? XD_DDA : XD_SVO_NAIVE;
    if (!o.shader.empty()) {
        traversal = -1;
        for (int t = 0; t < 5; ++t)
            if (o.shader == kShaders[t]) traversal = t;
        if (traversal < 0) throw Error("Invalid shader '" + o.shader + "'");
        const std::string need = traversal == XD_DDA ? "tiff" : "svo";
        if (need != type)
            throw Error("Shader '" + o.shader + "' is incompatible with model type '" + type + "' (requires '" + need +
                        "')");
    }
    log.log(std::string("Using shader '") + kShaders[traversal] + "'");

    uint64_t dim[3];
    std::map<xd_context*, xd_volume*> vols;
    if (type == "tiff") {
        GridData g = load_tiff(o.volume);
        std::copy(g.dim, g.dim + 3, dim);
        for (auto& out : outs)
            if (!vols.count(out.ctx)) check(xd_volume_from_grid(out.ctx, g.texels.data(), dim[0], dim[1], dim[2], &vols[out.ctx]));
    } else {
        SvoData s = load_svo(o.volume);
        dim[0] = dim[1] = dim[2] = s.side;
        for (auto& out : outs)
            if (!vols.count(out.ctx))
                check(xd_volume_from_svo(out.ctx, s.words.data(), s.node_count(), s.side, &vols[out.ctx]));
    }
    log.log(sformat("Model dimensions: %llux%llux%llu", (unsigned long long)dim[0], (unsigned long long)dim[1],
                    (unsigned long long)dim[2]));

    // display region = union of the outputs (RenderContext.cpp:43-57)
    int64_t x0 = outs[0].cfg.offset[0], y0 = outs[0].cfg.offset[1];
    int64_t x1 = x0 + outs[0].cfg.extent[0], y1 = y0 + outs[0].cfg.extent[1];
    for (auto& out : outs) {
        x0 = std::min<int64_t>(x0, out.cfg.offset[0]);
        y0 = std::min<int64_t>(y0, out.cfg.offset[1]);
        x1 = std::max<int64_t>(x1, out.cfg.offset[0] + (int64_t)out.cfg.extent[0]);
        y1 = std::max<int64_t>(y1, out.cfg.offset[1] + (int64_t)out.cfg.extent[1]);
    }
    const uint32_t dw = (uint32_t)(x1 - x0), dh = (uint32_t)(y1 - y0);
    log.log(sformat("Total resolution: %ux%u pixels", dw, dh));
    std::vector<xd_uniforms> unis(outs.size());
    for (size_t i = 0; i < outs.size(); ++i) {
        xd_uniforms& u = unis[i];
        memset(&u, 0, sizeof u);
        u.output_region = {{outs[i].cfg.offset[0], outs[i].cfg.offset[1]}, {outs[i].cfg.extent[0], outs[i].cfg.extent[1]}};
        u.display_region = {{(int32_t)x0, (int32_t)y0}, {dw, dh}};
        for (int k = 0; k < 3; ++k) {
            u.voxel_ratio[k] = o.ratio[k];
            u.model_dim[k] = (uint32_t)dim[k];
        }
        u.emission_coeff = (float)o.emission;
    }

    if (o.camera.empty() || o.camera == "orbit")
        throw Error("The orbit camera needs an interactive backend; pass --camera <file> with --headless");
    log.log("Reading camera transforms from '" + o.camera + "'");
    CameraScript script(o.camera);

    std::vector<RenderStats> all;
    std::vector<uint32_t> canvas;
    const auto start = std::chrono::steady_clock::now();
    auto fps_start = start;
    uint64_t frames = 0, total_frames = 0, saved = 0;
    log.log("Starting render loop...");
    for (;;) {
        ++frames;
        ++total_frames;
        // Renderer::render: one launch per output, camera translation divided
        // by the voxel ratio (Renderer.cpp:56-63)
        xd_camera cam{};
        for (int k = 0; k < 3; ++k) {
            cam.forward[k] = script.v[k];
            cam.up[k] = script.v[3 + k];
            cam.translation[k] = script.v[6 + k] / o.ratio[k];
        }
        RenderStats st;
        for (size_t i = 0; i < outs.size(); ++i) {
            double ms = 0;
            check(xd_render(outs[i].ctx, vols[outs[i].ctx], traversal, &cam, &unis[i], outs[i].frame, &ms));
            st.total_rays += (uint64_t)outs[i].cfg.extent[0] * outs[i].cfg.extent[1];
            st.outputs += 1;
            st.total += ms;
            st.max = std::max(st.max, ms);
            st.min = std::min(st.min, ms);
        }
        // HeadlessDisplay::swap_buffers / save (HeadlessDisplay.cpp:39-92)
        if (!o.discard) {
            const std::string path = format_output_path(o.output, saved);
            log.log(sformat("Saving frame %llu...", (unsigned long long)saved));
            canvas.assign((size_t)dw * dh, 0xFF000000u);
            for (size_t i = 0; i < outs.size(); ++i) {
                const size_t off = (size_t)(outs[i].cfg.offset[1] - y0) * dw + (size_t)(outs[i].cfg.offset[0] - x0);
                check(xd_frame_download(outs[i].frame, canvas.data() + off, dw));
            }
            log.log("Compressing...");
            try {
                save_png(path, canvas.data(), dw, dh);
                log.log("Saved output to '" + path + "'");
            } catch (const Error& e) {
                log.log(std::string("Error saving output: ") + e.what());
            }
        }
        ++saved;
        all.push_back(st);
        if (total_frames % (uint64_t)o.repeat == 0 && script.update()) break;
        const auto now = std::chrono::steady_clock::now();
        const double diff = std::chrono::duration<double>(now - fps_start).count();
        if (diff > 5.0) {
            log.log("FPS: " + fmt_double((double)frames / diff));
            frames = 0;
            fps_start = now;
        }
    }
    const double total_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    uint64_t rays = 0;
    double time = 0;
    for (const auto& f : all) {
        rays += f.total_rays;
        time += f.total;
    }
    const double fps = (double)all.size() / total_s;
    log.log(sformat("total rays: %llu, total render time: ", (unsigned long long)rays) + fmt_double(time) +
            "ms, mray/s: " + fmt_double((double)rays / (time * 1000.0)));
    log.log(sformat("frames: %zu, fps: ", all.size()) + fmt_double(fps) + ", total time: " + fmt_double(total_s) + "s");
    if (!o.stats.empty()) {
        save_stats(o.stats, all, fps);
        log.log("Saved stats to '" + o.stats + "'");
    }
}

void render(const std::vector<std::string>& args) {
    RenderOptions o;
    try {
        o = parse_render_args(args);
    } catch (const Error& e) {
        printf("Error: %s\n", e.what());
        return;
    }
    Logger log;
    if (!o.quiet) log.add_console();
    if (!o.log_output.empty()) log.add_file(o.log_output);

    std::vector<Output> outs;
    std::map<uint32_t, xd_context*> ctxs;
    auto close_all = [&]() {
        for (auto& kv : ctxs) xd_close(kv.second);  // frees volumes and frames too
        ctxs.clear();
    };
    try {
        if (o.xorg || !o.direct.empty())
            throw Error(std::string("the ") + (o.xorg ? "xorg" : "direct") +
                        " backend is not part of this build; use --headless");
        log.log("Using headless presenting backend");
        std::ifstream in(o.headless);
        if (!in) throw Error("Failed to open config file '" + o.headless + "'");
        std::vector<HeadlessDevice> devs;
        try {
            devs = parse_headless_config(in);
        } catch (const Error& e) {
            throw Error("Failed to read config file '" + o.headless + "': " + e.what());
        }
        int count = 0;
        check(xd_device_count(&count));
        for (const auto& d : devs) {
            if (d.vkindex >= (uint32_t)count)
                throw Error(sformat("no GPU with index %u (%d detected)", d.vkindex, count));
            if (d.extent[0] == 0 || d.extent[1] == 0) throw Error("a device region has an empty extent");
            if (!ctxs.count(d.vkindex)) check(xd_open((int)d.vkindex, &ctxs[d.vkindex]));
            Output out;
            out.cfg = d;
            out.ctx = ctxs[d.vkindex];
            check(xd_frame_create(out.ctx, d.extent[0], d.extent[1], &out.frame));
            outs.push_back(out);
        }
    } catch (const Error& e) {
        printf("Error: Failed to initialize backend: %s\n", e.what());
        close_all();
        return;
    }
    try {
        main_loop(log, o, outs);
    } catch (const Error& e) {
        printf("Error: %s\n", e.what());
    }
    close_all();
}

}  // namespace

int main(int argc, const char* argv[]) {
    if (argc < 2) {
        help(argv[0], {});
        return 0;
    }
    const std::vector<std::string> args(argv + 2, argv + argc);
    const std::string sub = argv[1];
    if (sub == "help")
        help(argv[0], args);
    else if (sub == "sysinfo")
        sysinfo();
    else if (sub == "render")
        render(args);
    else if (sub == "convert")
        convert(args);
    else
        printf("Error: Invalid subcommand '%s', see '%s help'\n", sub.c_str(), argv[0]);
    fflush(stdout);
    return 0;
}
