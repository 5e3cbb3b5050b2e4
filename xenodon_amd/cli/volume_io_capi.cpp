// C entry points of the CLI's file formats (libxd_io.so), for scripts and
// tests: every function returns 0 on success, else writes the error message
// to err (err_len bytes) and returns 1.
#include <cstdlib>
#include <cstring>

#include "cli_util.hpp"
#include "volume_io.hpp"

using namespace xdcli;

namespace {
int fail(const std::exception& e, char* err, size_t n) {
    if (err && n) {
        strncpy(err, e.what(), n - 1);
        err[n - 1] = '\0';
    }
    return 1;
}
}  // namespace

extern "C" {

__attribute__((visibility("default"))) int xdio_load_tiff(const char* path, uint64_t dims[3], uint32_t** texels,
                                                          char* err, size_t err_len) {
    try {
        GridData g = load_tiff(path, false);
        dims[0] = g.dim[0], dims[1] = g.dim[1], dims[2] = g.dim[2];
        *texels = static_cast<uint32_t*>(malloc(g.texels.size() * 4 + 4));
        memcpy(*texels, g.texels.data(), g.texels.size() * 4);
        return 0;
    } catch (const std::exception& e) {
        return fail(e, err, err_len);
    }
}

__attribute__((visibility("default"))) int xdio_save_tiff(const char* path, const uint64_t dims[3],
                                                          const uint32_t* texels, char* err, size_t err_len) {
    try {
        GridData g;
        g.dim[0] = dims[0], g.dim[1] = dims[1], g.dim[2] = dims[2];
        g.texels.assign(texels, texels + dims[0] * dims[1] * dims[2]);
        save_tiff(path, g);
        return 0;
    } catch (const std::exception& e) {
        return fail(e, err, err_len);
    }
}

__attribute__((visibility("default"))) int xdio_load_svo(const char* path, uint64_t* side, uint64_t* node_count,
                                                         uint32_t** words, char* err, size_t err_len) {
    try {
        SvoData s = load_svo(path);
        *side = s.side;
        *node_count = s.node_count();
        *words = static_cast<uint32_t*>(malloc(s.words.size() * 4 + 4));
        memcpy(*words, s.words.data(), s.words.size() * 4);
        return 0;
    } catch (const std::exception& e) {
        return fail(e, err, err_len);
    }
}

__attribute__((visibility("default"))) int xdio_save_svo(const char* path, uint64_t side, uint64_t node_count,
                                                         const uint32_t* words, char* err, size_t err_len) {
    try {
        SvoData s;
        s.side = side;
        s.words.assign(words, words + node_count * 10);
        save_svo(path, s);
        return 0;
    } catch (const std::exception& e) {
        return fail(e, err, err_len);
    }
}

__attribute__((visibility("default"))) int xdio_save_png(const char* path, const uint32_t* rgba, uint32_t w,
                                                         uint32_t h, char* err, size_t err_len) {
    try {
        save_png(path, rgba, w, h);
        return 0;
    } catch (const std::exception& e) {
        return fail(e, err, err_len);
    }
}

__attribute__((visibility("default"))) int xdio_format_output_path(const char* fmt, uint64_t frame, char* out,
                                                                   size_t out_len, char* err, size_t err_len) {
    try {
        const std::string s = format_output_path(fmt, frame);
        if (s.size() + 1 > out_len) throw Error("output buffer too small");
        memcpy(out, s.c_str(), s.size() + 1);
        return 0;
    } catch (const std::exception& e) {
        return fail(e, err, err_len);
    }
}

__attribute__((visibility("default"))) void xdio_free(void* p) { free(p); }
}
