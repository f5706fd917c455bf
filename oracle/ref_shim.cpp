// oracle/ref_shim.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C wrapper around the reference quicktex C++ codecs, compiled together with
// the reference's own core translation units (see oracle/Makefile, target
// `ref`) into oracle/_ref/libqtxref.so.  Only tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg load it, as a checker or as the timed CPU
// baseline; nothing in quicktex_amd/ links or calls it.
//
// Entry points mirror the Python bindings the reference exposes
// (quicktex/s3tc/*/_bindings.cpp) plus the advanced BC1 properties
// (quicktex/s3tc/bc1/_bindings.cpp:134-164), so any qtx_bc1_params value can
// be replayed on the reference encoder.
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>

#include "../include/qtx.h"
#include "Texture.h"
#include "s3tc/bc1/BC1Decoder.h"
#include "s3tc/bc1/BC1Encoder.h"
#include "s3tc/bc3/BC3Decoder.h"
#include "s3tc/bc3/BC3Encoder.h"
#include "s3tc/bc4/BC4Decoder.h"
#include "s3tc/bc4/BC4Encoder.h"
#include "s3tc/bc5/BC5Decoder.h"
#include "s3tc/bc5/BC5Encoder.h"
#include "s3tc/interpolator/Interpolator.h"

using namespace quicktex;
using namespace quicktex::s3tc;

namespace {
thread_local std::string g_err;
std::mutex g_mu;
std::map<std::pair<int, int>, std::shared_ptr<BC1Encoder>> g_bc1;
std::map<int, std::shared_ptr<BC3Encoder>> g_bc3;

std::shared_ptr<Interpolator> make_interp(int t) {
    return std::shared_ptr<Interpolator>(Interpolator::MakeInterpolator(static_cast<Interpolator::Type>(t)));
}

void apply(BC1Encoder &e, const qtx_bc1_params *p) {
    e.SetErrorMode(static_cast<BC1Encoder::ErrorMode>(p->error_mode));
    e.SetEndpointMode(static_cast<BC1Encoder::EndpointMode>(p->endpoint_mode));
    e.two_ls_passes = p->two_ls_passes != 0;
    e.two_ep_passes = p->two_ep_passes != 0;
    e.two_cf_passes = p->two_cf_passes != 0;
    e.exhaustive = p->exhaustive != 0;
    e.SetPowerIterations(p->power_iterations);
    e.SetSearchRounds(p->search_rounds);
    e.SetOrderings4(p->orderings4);
    e.SetOrderings3(p->orderings3);
}

RawTexture to_raw(const uint8_t *rgba, int w, int h) {
    RawTexture t(w, h);
    std::memcpy(t.Data(), rgba, t.NBytes());
    return t;
}

template <class B> BlockTexture<B> to_blocks(const uint8_t *data, int w, int h) {
    BlockTexture<B> t(w, h);
    std::memcpy(t.Data(), data, t.NBytes());
    return t;
}

template <class T> void copy_out(const T &t, uint8_t *out) { std::memcpy(out, t.Data(), t.NBytes()); }

template <class F> int guarded(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
        return 1;
    }
}
}  // namespace

extern "C" {

const char *qtxr_last_error(void) { return g_err.c_str(); }

// Parameters as the reference's constructor + SetLevel leave them.
int qtxr_bc1_params_init(qtx_bc1_params *out, uint32_t level, int32_t color_mode, int32_t interp) {
    return guarded([&] {
        BC1Encoder e(level, static_cast<BC1Encoder::ColorMode>(color_mode), make_interp(interp));
        out->interpolator = interp;
        out->color_mode = color_mode;
        out->error_mode = static_cast<int32_t>(e.GetErrorMode());
        out->endpoint_mode = static_cast<int32_t>(e.GetEndpointMode());
        out->two_ls_passes = e.two_ls_passes;
        out->two_ep_passes = e.two_ep_passes;
        out->two_cf_passes = e.two_cf_passes;
        out->exhaustive = e.exhaustive;
        out->power_iterations = e.GetPowerIterations();
        out->search_rounds = e.GetSearchRounds();
        out->orderings4 = e.GetOrderings4();
        out->orderings3 = e.GetOrderings3();
    });
}

// `in` replayed through the reference's setters (apply()), read back through its
// getters: the values EncodeBlock then sees (BC1Encoder.cpp:251-258 clamps).
int qtxr_bc1_params_apply(const qtx_bc1_params *in, qtx_bc1_params *out) {
    return guarded([&] {
        BC1Encoder e(5, static_cast<BC1Encoder::ColorMode>(in->color_mode), make_interp(in->interpolator));
        apply(e, in);
        *out = *in;
        out->error_mode = static_cast<int32_t>(e.GetErrorMode());
        out->endpoint_mode = static_cast<int32_t>(e.GetEndpointMode());
        out->two_ls_passes = e.two_ls_passes;
        out->two_ep_passes = e.two_ep_passes;
        out->two_cf_passes = e.two_cf_passes;
        out->exhaustive = e.exhaustive;
        out->power_iterations = e.GetPowerIterations();
        out->search_rounds = e.GetSearchRounds();
        out->orderings4 = e.GetOrderings4();
        out->orderings3 = e.GetOrderings3();
    });
}

int qtxr_bc1_encode(const qtx_bc1_params *p, const uint8_t *rgba, int32_t w, int32_t h, uint8_t *out) {
    return guarded([&] {
        std::shared_ptr<BC1Encoder> enc;
        {
            std::lock_guard<std::mutex> lk(g_mu);
            auto key = std::make_pair(p->color_mode, p->interpolator);
            auto it = g_bc1.find(key);
            if (it == g_bc1.end()) {
                enc = std::make_shared<BC1Encoder>(5, static_cast<BC1Encoder::ColorMode>(p->color_mode),
                                                   make_interp(p->interpolator));
                g_bc1[key] = enc;
            } else {
                enc = it->second;
            }
        }
        apply(*enc, p);
        copy_out(enc->Encode(to_raw(rgba, w, h)), out);
    });
}

int qtxr_bc3_encode(const qtx_bc1_params *p, const uint8_t *rgba, int32_t w, int32_t h, uint8_t *out) {
    return guarded([&] {
        std::shared_ptr<BC3Encoder> enc;
        {
            std::lock_guard<std::mutex> lk(g_mu);
            auto it = g_bc3.find(p->interpolator);
            if (it == g_bc3.end()) {
                enc = std::make_shared<BC3Encoder>(5, make_interp(p->interpolator));
                g_bc3[p->interpolator] = enc;
            } else {
                enc = it->second;
            }
        }
        apply(*enc->GetBC1Encoder(), p);
        copy_out(enc->Encode(to_raw(rgba, w, h)), out);
    });
}

int qtxr_bc4_encode(uint32_t ch, const uint8_t *rgba, int32_t w, int32_t h, uint8_t *out) {
    return guarded([&] { copy_out(BC4Encoder(static_cast<uint8_t>(ch)).Encode(to_raw(rgba, w, h)), out); });
}

int qtxr_bc5_encode(uint32_t c0, uint32_t c1, const uint8_t *rgba, int32_t w, int32_t h, uint8_t *out) {
    return guarded([&] {
        copy_out(BC5Encoder(static_cast<uint8_t>(c0), static_cast<uint8_t>(c1)).Encode(to_raw(rgba, w, h)), out);
    });
}

int qtxr_bc1_decode(int32_t interp, int32_t write_alpha, const uint8_t *blocks, int32_t w, int32_t h, uint8_t *out) {
    return guarded([&] {
        BC1Decoder d(write_alpha != 0, make_interp(interp));
        copy_out(d.Decode(to_blocks<BC1Block>(blocks, w, h)), out);
    });
}

int qtxr_bc3_decode(int32_t interp, const uint8_t *blocks, int32_t w, int32_t h, uint8_t *out) {
    return guarded([&] {
        BC3Decoder d(make_interp(interp));
        copy_out(d.Decode(to_blocks<BC3Block>(blocks, w, h)), out);
    });
}

int qtxr_bc4_decode(uint32_t ch, const uint8_t *blocks, int32_t w, int32_t h, uint8_t *out) {
    return guarded([&] { copy_out(BC4Decoder(static_cast<uint8_t>(ch)).Decode(to_blocks<BC4Block>(blocks, w, h)), out); });
}

int qtxr_bc5_decode(uint32_t c0, uint32_t c1, const uint8_t *blocks, int32_t w, int32_t h, uint8_t *out) {
    return guarded([&] {
        BC5Decoder d(static_cast<uint8_t>(c0), static_cast<uint8_t>(c1));
        copy_out(d.Decode(to_blocks<BC5Block>(blocks, w, h)), out);
    });
}

}  // extern "C"
