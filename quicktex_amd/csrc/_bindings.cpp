// _bindings.cpp -- the pybind11 module `_quicktex` of quicktex_amd.
//
// Same Python surface as the reference extension (quicktex/_bindings.cpp,
// quicktex/s3tc/_bindings.cpp, quicktex/s3tc/{bc1,bc3,bc4,bc5,interpolator}/_bindings.cpp):
// module `_quicktex` with `Texture`, `RawTexture` and submodule `_s3tc` holding
// `_interpolator`, `_bc1`, `_bc4`, `_bc3`, `_bc5`.  Containers live in host
// memory here; every encode and decode is one call into the C ABI
// (include/qtx.h), which runs the gfx950 kernels.  There is no CPU codec.
//
// Exceptions follow the reference's pybind11 translation: bad values ->
// ValueError, indices -> IndexError, buffer-format problems -> RuntimeError.
// The GIL is released while the device works.
#include <pybind11/operators.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "qtx.h"

namespace py = pybind11;
using namespace pybind11::literals;

namespace qb {

// ---------------------------------------------------------------------------
// Colours cross the boundary as (r, g, b[, a]) tuples (quicktex/_bindings.h:34-77).
struct Color {
    uint8_t r = 0, g = 0, b = 0, a = 255;
};

}  // namespace qb

namespace pybind11::detail {
template <> struct type_caster<qb::Color> {
    PYBIND11_TYPE_CASTER(qb::Color, const_name("Color"));
    bool load(handle src, bool) {
        if (!PySequence_Check(src.ptr()) || PyUnicode_Check(src.ptr())) return false;
        sequence seq = reinterpret_borrow<sequence>(src);
        const size_t n = seq.size();
        if (n < 3 || n > 4) return false;
        uint8_t ch[4] = {0, 0, 0, 255};
        for (size_t i = 0; i < n; i++) {
            object item = seq[i];
            if (!PyNumber_Check(item.ptr())) return false;
            const long v = PyLong_AsLong(item.ptr());
            if (PyErr_Occurred()) {
                PyErr_Clear();
                return false;
            }
            if (v < 0 || v > 255) return false;
            ch[i] = (uint8_t)v;
        }
        value.r = ch[0];
        value.g = ch[1];
        value.b = ch[2];
        value.a = ch[3];
        return true;
    }
    static handle cast(const qb::Color &c, return_value_policy, handle) {
        return make_tuple((int)c.r, (int)c.g, (int)c.b, (int)c.a).release();
    }
};
}  // namespace pybind11::detail

namespace qb {

// ---------------------------------------------------------------------------
// C-ABI plumbing: one context per process on device QTX_DEVICE (default 0).
[[noreturn]] void raise_status(qtx_status st) {
    const std::string msg = qtx_last_error();
    if (st == QTX_EINVAL) throw std::invalid_argument(msg);
    if (st == QTX_ERANGE) throw std::out_of_range(msg);
    throw std::runtime_error(msg);
}
inline void check(qtx_status st) {
    if (st != QTX_OK) raise_status(st);
}

qtx_ctx *context() {
    static std::once_flag once;
    static qtx_ctx *ctx = nullptr;
    static qtx_status st = QTX_OK;
    std::call_once(once, [] {
        const char *env = std::getenv("QTX_DEVICE");
        st = qtx_ctx_create(env ? std::atoi(env) : 0, &ctx);
    });
    if (!ctx) raise_status(st);
    return ctx;
}

// quicktex/_bindings.h:129-133
int py_index(int val, int size, const std::string &name = "index") {
    if (val < -size || val >= size) throw std::out_of_range(name + " value out of range");
    return val < 0 ? size + val : val;
}

uint16_t pack565(const Color &c) {  // Color::Pack565 (quicktex/Color.cpp:145)
    auto s5 = [](int v) { int t = v * 31 + 128; return (t + (t >> 8)) >> 8; };
    auto s6 = [](int v) { int t = v * 63 + 128; return (t + (t >> 8)) >> 8; };
    return (uint16_t)((s5(c.r) << 11) | (s6(c.g) << 5) | s5(c.b));
}
Color unpack565(uint16_t p) {  // Color::Unpack565 (quicktex/Color.cpp:155-161)
    const int r = (p >> 11) & 31, g = (p >> 5) & 63, b = p & 31;
    return Color{(uint8_t)((r << 3) | (r >> 2)), (uint8_t)((g << 2) | (g >> 4)), (uint8_t)((b << 3) | (b >> 2)), 255};
}

using Sel4x4 = std::array<std::array<uint8_t, 4>, 4>;

// ---------------------------------------------------------------------------
// Blocks (quicktex/s3tc/*/BC*Block.h); layouts in the comments of include/qtx.h.
struct BC1Block {
    static constexpr int Width = 4, Height = 4;
    uint8_t b[8] = {0};
    BC1Block() = default;
    BC1Block(Color c0, Color c1, const Sel4x4 &s) {
        set_raw(0, pack565(c0));
        set_raw(1, pack565(c1));
        set_selectors(s);
    }
    uint16_t raw(int i) const { return (uint16_t)(b[2 * i] | (b[2 * i + 1] << 8)); }
    void set_raw(int i, uint16_t v) {
        b[2 * i] = (uint8_t)v;
        b[2 * i + 1] = (uint8_t)(v >> 8);
    }
    std::pair<Color, Color> endpoints() const { return {unpack565(raw(0)), unpack565(raw(1))}; }
    void set_endpoints(std::pair<Color, Color> e) {
        set_raw(0, pack565(e.first));
        set_raw(1, pack565(e.second));
    }
    Sel4x4 selectors() const {
        Sel4x4 s;
        for (int y = 0; y < 4; y++)
            for (int x = 0; x < 4; x++) s[y][x] = (b[4 + y] >> (2 * x)) & 3;
        return s;
    }
    void set_selectors(const Sel4x4 &s) {  // BC1Block.cpp:33-41
        for (auto &row : s)
            for (auto v : row)
                if (v > 3) throw std::invalid_argument("Selector value out of bounds.");
        for (int y = 0; y < 4; y++) {
            uint8_t r = 0;
            for (int x = 0; x < 4; x++) r |= (uint8_t)(s[y][x] << (2 * x));
            b[4 + y] = r;
        }
    }
    bool is_3color() const { return raw(0) <= raw(1); }
    bool operator==(const BC1Block &o) const { return std::memcmp(b, o.b, 8) == 0; }
};

struct BC4Block {
    static constexpr int Width = 4, Height = 4;
    uint8_t b[8] = {0};
    BC4Block() = default;
    BC4Block(uint8_t a0, uint8_t a1, const Sel4x4 &s) {
        b[0] = a0;
        b[1] = a1;
        set_selectors(s);
    }
    std::pair<uint8_t, uint8_t> endpoints() const { return {b[0], b[1]}; }
    void set_endpoints(std::pair<uint8_t, uint8_t> e) {
        b[0] = e.first;
        b[1] = e.second;
    }
    uint64_t bits() const {
        uint64_t v = 0;
        for (int k = 0; k < 6; k++) v |= (uint64_t)b[2 + k] << (8 * k);
        return v;
    }
    Sel4x4 selectors() const {
        const uint64_t v = bits();
        Sel4x4 s;
        for (int i = 0; i < 16; i++) s[i / 4][i % 4] = (uint8_t)((v >> (3 * i)) & 7);
        return s;
    }
    void set_selectors(const Sel4x4 &s) {  // BC4Block.cpp:34-42
        for (auto &row : s)
            for (auto v : row)
                if (v > 7) throw std::invalid_argument("Selector value out of bounds.");
        uint64_t v = 0;
        for (int i = 0; i < 16; i++) v |= (uint64_t)s[i / 4][i % 4] << (3 * i);
        for (int k = 0; k < 6; k++) b[2 + k] = (uint8_t)(v >> (8 * k));
    }
    bool is_6value() const { return b[0] <= b[1]; }
    std::array<uint8_t, 8> values() const {  // BC4Block.cpp:44-64
        const int a0 = b[0], a1 = b[1];
        if (a0 <= a1)
            return {(uint8_t)a0, (uint8_t)a1, (uint8_t)((a0 * 4 + a1) / 5), (uint8_t)((a0 * 3 + a1 * 2) / 5),
                    (uint8_t)((a0 * 2 + a1 * 3) / 5), (uint8_t)((a0 + a1 * 4) / 5), 0, 255};
        return {(uint8_t)a0, (uint8_t)a1, (uint8_t)((a0 * 6 + a1) / 7), (uint8_t)((a0 * 5 + a1 * 2) / 7),
                (uint8_t)((a0 * 4 + a1 * 3) / 7), (uint8_t)((a0 * 3 + a1 * 4) / 7), (uint8_t)((a0 * 2 + a1 * 5) / 7),
                (uint8_t)((a0 + a1 * 6) / 7)};
    }
    bool operator==(const BC4Block &o) const { return std::memcmp(b, o.b, 8) == 0; }
};

struct BC3Block {  // alpha block first (BC3Block.h:35-36)
    static constexpr int Width = 4, Height = 4;
    BC4Block alpha_block;
    BC1Block color_block;
    BC3Block() = default;
    BC3Block(const BC4Block &a, const BC1Block &c) : alpha_block(a), color_block(c) {}
    std::pair<BC4Block, BC1Block> blocks() const { return {alpha_block, color_block}; }
    void set_blocks(const std::pair<BC4Block, BC1Block> &p) {
        alpha_block = p.first;
        color_block = p.second;
    }
    bool operator==(const BC3Block &o) const { return alpha_block == o.alpha_block && color_block == o.color_block; }
};

struct BC5Block {
    static constexpr int Width = 4, Height = 4;
    BC4Block chan0_block, chan1_block;
    BC5Block() = default;
    BC5Block(const BC4Block &a, const BC4Block &b) : chan0_block(a), chan1_block(b) {}
    std::pair<BC4Block, BC4Block> blocks() const { return {chan0_block, chan1_block}; }
    void set_blocks(const std::pair<BC4Block, BC4Block> &p) {
        chan0_block = p.first;
        chan1_block = p.second;
    }
    bool operator==(const BC5Block &o) const { return chan0_block == o.chan0_block && chan1_block == o.chan1_block; }
};
static_assert(sizeof(BC1Block) == 8 && sizeof(BC4Block) == 8 && sizeof(BC3Block) == 16 && sizeof(BC5Block) == 16);

// ---------------------------------------------------------------------------
// Textures (quicktex/Texture.h)
struct Texture {
    Texture(int w, int h) : width(w), height(h) {
        if (w <= 0) throw std::invalid_argument("Texture width must be greater than 0");
        if (h <= 0) throw std::invalid_argument("Texture height must be greater than 0");
    }
    virtual ~Texture() = default;
    virtual size_t nbytes() const = 0;
    virtual uint8_t *data() = 0;
    const uint8_t *cdata() const { return const_cast<Texture *>(this)->data(); }
    std::tuple<int, int> size() const { return {width, height}; }
    int width, height;
};

struct RawTexture : Texture {
    RawTexture(int w, int h) : Texture(w, h), px((size_t)w * (size_t)h * 4, 0) {
        for (size_t i = 3; i < px.size(); i += 4) px[i] = 255;  // Color() = (0, 0, 0, 255)
    }
    size_t nbytes() const override { return px.size(); }
    uint8_t *data() override { return px.data(); }
    Color get(int x, int y) const {
        const uint8_t *p = &px[4 * ((size_t)y * width + x)];
        return Color{p[0], p[1], p[2], p[3]};
    }
    void set(int x, int y, Color c) {
        uint8_t *p = &px[4 * ((size_t)y * width + x)];
        p[0] = c.r; p[1] = c.g; p[2] = c.b; p[3] = c.a;
    }
    std::vector<uint8_t> px;
};

template <class B> struct BlockTexture : Texture {
    BlockTexture(int w, int h) : Texture(w, h), bw((w + 3) / 4), bh((h + 3) / 4), blocks((size_t)bw * bh) {}
    size_t nbytes() const override { return blocks.size() * sizeof(B); }
    uint8_t *data() override { return reinterpret_cast<uint8_t *>(blocks.data()); }
    std::tuple<int, int> size_blocks() const { return {bw, bh}; }
    int bw, bh;
    std::vector<B> blocks;
};

// BufferToTexture / BufferToPOD (quicktex/_bindings.h:89-127)
py::buffer_info request_bytes(const py::buffer &buf, size_t need) {
    py::buffer_info info = buf.request(false);
    if (info.format != py::format_descriptor<uint8_t>::format())
        throw std::runtime_error("Incompatible format in python buffer: expected a byte array.");
    if (info.ndim != 1) throw std::runtime_error("Incompatible format in python buffer: Incorrect number of dimensions.");
    if (info.shape[0] < (py::ssize_t)need)
        throw std::runtime_error("Incompatible format in python buffer: 1-D buffer has incorrect length.");
    if (info.strides[0] != 1) throw std::runtime_error("Incompatible format in python buffer: 1-D buffer is not contiguous.");
    return info;
}
template <class T> std::unique_ptr<T> texture_from_bytes(py::buffer buf, int w, int h) {
    auto t = std::make_unique<T>(w, h);
    py::buffer_info info = request_bytes(buf, t->nbytes());
    std::memcpy(t->data(), info.ptr, t->nbytes());
    return t;
}
template <class B> B block_from_bytes(py::buffer buf) {
    py::buffer_info info = request_bytes(buf, sizeof(B));
    B b;
    std::memcpy(&b, info.ptr, sizeof(B));
    return b;
}

// ---------------------------------------------------------------------------
// Interpolators: tags only; the arithmetic lives in the kernels.
struct Interpolator {
    virtual ~Interpolator() = default;
    virtual int type() const { return QTX_INTERP_IDEAL; }
};
struct InterpolatorRound : Interpolator {
    int type() const override { return QTX_INTERP_IDEAL_ROUND; }
};
struct InterpolatorNvidia : Interpolator {
    int type() const override { return QTX_INTERP_NVIDIA; }
};
struct InterpolatorAMD : Interpolator {
    int type() const override { return QTX_INTERP_AMD; }
};
using InterpPtr = std::shared_ptr<Interpolator>;
InterpPtr interp_ptr(InterpPtr p) { return p ? p : std::make_shared<Interpolator>(); }

enum class ColorMode { FourColor = QTX_COLOR_FOUR, ThreeColor = QTX_COLOR_THREE, ThreeColorBlack = QTX_COLOR_THREE_BLACK };
enum class ErrorMode { None = 0, Faster = 1, Check2 = 2, Full = 3 };
enum class EndpointMode { LeastSquares = 0, BoundingBox = 1, BoundingBoxInt = 2, PCA = 3 };

// ---------------------------------------------------------------------------
// Encoders and decoders
struct BC1Encoder {
    BC1Encoder(unsigned level, ColorMode mode, InterpPtr interp) : interpolator(interp_ptr(interp)), color_mode(mode) {
        check(qtx_bc1_params_init(&p, level, (int32_t)mode, interpolator->type()));
    }
    void set_level(unsigned level) { check(qtx_bc1_params_init(&p, level, (int32_t)color_mode, interpolator->type())); }
    std::unique_ptr<BlockTexture<BC1Block>> encode(const RawTexture &t) const {
        auto out = std::make_unique<BlockTexture<BC1Block>>(t.width, t.height);
        qtx_ctx *ctx = context();
        qtx_status st;
        {
            py::gil_scoped_release nogil;
            st = qtx_bc1_encode(ctx, &p, t.cdata(), t.width, t.height, out->data(), out->nbytes());
        }
        check(st);
        return out;
    }
    std::tuple<unsigned, unsigned> orderings() const { return {p.orderings4, p.orderings3}; }
    void set_orderings(std::tuple<unsigned, unsigned> o) {
        p.orderings4 = std::get<0>(o);
        p.orderings3 = std::get<1>(o);
        check(qtx_bc1_params_clamp(&p));
    }
    void set_power_iterations(unsigned n) {
        p.power_iterations = n;
        check(qtx_bc1_params_clamp(&p));
    }
    InterpPtr interpolator;
    ColorMode color_mode;
    qtx_bc1_params p{};
};

struct BC1Decoder {
    BC1Decoder(bool wa, InterpPtr i) : write_alpha(wa), interpolator(interp_ptr(i)) {}
    std::unique_ptr<RawTexture> decode(const BlockTexture<BC1Block> &t) const {
        auto out = std::make_unique<RawTexture>(t.width, t.height);
        qtx_ctx *ctx = context();
        qtx_status st;
        {
            py::gil_scoped_release nogil;
            st = qtx_bc1_decode(ctx, interpolator->type(), write_alpha, t.cdata(), t.width, t.height, out->data(),
                                out->nbytes());
        }
        check(st);
        return out;
    }
    bool write_alpha;
    InterpPtr interpolator;
};

struct BC4Encoder {
    explicit BC4Encoder(uint8_t ch) : channel(ch) {
        if (ch >= 4) throw std::invalid_argument("Channel out of range");
    }
    std::unique_ptr<BlockTexture<BC4Block>> encode(const RawTexture &t) const {
        auto out = std::make_unique<BlockTexture<BC4Block>>(t.width, t.height);
        qtx_ctx *ctx = context();
        qtx_status st;
        {
            py::gil_scoped_release nogil;
            st = qtx_bc4_encode(ctx, channel, t.cdata(), t.width, t.height, out->data(), out->nbytes());
        }
        check(st);
        return out;
    }
    uint8_t channel;
};

struct BC4Decoder {
    explicit BC4Decoder(uint8_t ch) : channel(ch) {
        if (ch >= 4) throw std::invalid_argument("Channel out of range");
    }
    std::unique_ptr<RawTexture> decode(const BlockTexture<BC4Block> &t) const {
        auto out = std::make_unique<RawTexture>(t.width, t.height);
        qtx_ctx *ctx = context();
        qtx_status st;
        {
            py::gil_scoped_release nogil;
            st = qtx_bc4_decode(ctx, channel, t.cdata(), t.width, t.height, out->data(), out->nbytes());
        }
        check(st);
        return out;
    }
    uint8_t channel;
};

struct BC3Encoder {
    BC3Encoder(unsigned level, InterpPtr i)
        : bc1(std::make_shared<BC1Encoder>(level, ColorMode::FourColor, i)), bc4(std::make_shared<BC4Encoder>(3)) {}
    std::unique_ptr<BlockTexture<BC3Block>> encode(const RawTexture &t) const {
        auto out = std::make_unique<BlockTexture<BC3Block>>(t.width, t.height);
        qtx_ctx *ctx = context();
        qtx_status st;
        {
            py::gil_scoped_release nogil;
            st = qtx_bc3_encode(ctx, &bc1->p, t.cdata(), t.width, t.height, out->data(), out->nbytes());
        }
        check(st);
        return out;
    }
    std::shared_ptr<BC1Encoder> bc1;
    std::shared_ptr<BC4Encoder> bc4;
};

struct BC3Decoder {
    explicit BC3Decoder(InterpPtr i)
        : bc1(std::make_shared<BC1Decoder>(false, i)), bc4(std::make_shared<BC4Decoder>(3)) {}
    std::unique_ptr<RawTexture> decode(const BlockTexture<BC3Block> &t) const {
        auto out = std::make_unique<RawTexture>(t.width, t.height);
        qtx_ctx *ctx = context();
        qtx_status st;
        {
            py::gil_scoped_release nogil;
            st = qtx_bc3_decode(ctx, bc1->interpolator->type(), t.cdata(), t.width, t.height, out->data(),
                                out->nbytes());
        }
        check(st);
        return out;
    }
    std::shared_ptr<BC1Decoder> bc1;
    std::shared_ptr<BC4Decoder> bc4;
};

struct BC5Encoder {
    BC5Encoder(uint8_t c0, uint8_t c1) : e0(std::make_shared<BC4Encoder>(c0)), e1(std::make_shared<BC4Encoder>(c1)) {}
    std::unique_ptr<BlockTexture<BC5Block>> encode(const RawTexture &t) const {
        auto out = std::make_unique<BlockTexture<BC5Block>>(t.width, t.height);
        qtx_ctx *ctx = context();
        qtx_status st;
        {
            py::gil_scoped_release nogil;
            st = qtx_bc5_encode(ctx, e0->channel, e1->channel, t.cdata(), t.width, t.height, out->data(), out->nbytes());
        }
        check(st);
        return out;
    }
    std::shared_ptr<BC4Encoder> e0, e1;
};

struct BC5Decoder {
    BC5Decoder(uint8_t c0, uint8_t c1) : d0(std::make_shared<BC4Decoder>(c0)), d1(std::make_shared<BC4Decoder>(c1)) {}
    std::unique_ptr<RawTexture> decode(const BlockTexture<BC5Block> &t) const {
        auto out = std::make_unique<RawTexture>(t.width, t.height);
        qtx_ctx *ctx = context();
        qtx_status st;
        {
            py::gil_scoped_release nogil;
            st = qtx_bc5_decode(ctx, d0->channel, d1->channel, t.cdata(), t.width, t.height, out->data(), out->nbytes());
        }
        check(st);
        return out;
    }
    std::shared_ptr<BC4Decoder> d0, d1;
};

// ---------------------------------------------------------------------------
// Binding helpers (quicktex/_bindings.h:135-234)
template <class B> py::class_<B> bind_block(py::module_ &m, const char *name) {
    py::class_<B> c(m, name, py::buffer_protocol());
    c.def_static("frombytes", &block_from_bytes<B>, "data"_a);
    c.def_property_readonly_static("width", [](py::object) { return B::Width; });
    c.def_property_readonly_static("height", [](py::object) { return B::Height; });
    c.def_property_readonly_static("size", [](py::object) { return std::make_tuple(B::Width, B::Height); });
    c.def_property_readonly_static("nbytes", [](py::object) { return sizeof(B); });
    c.def(py::self == py::self);
    c.def_buffer([](B &b) { return py::buffer_info(reinterpret_cast<uint8_t *>(&b), (py::ssize_t)sizeof(B)); });
    c.def("tobytes", [](const B &b) { return py::bytes(reinterpret_cast<const char *>(&b), sizeof(B)); });
    return c;
}

template <class B> void bind_block_texture(py::module_ &m, const char *name) {
    using T = BlockTexture<B>;
    py::class_<T, Texture> c(m, name);
    c.def(py::init<int, int>(), "width"_a, "height"_a);
    c.def_static("from_bytes", &texture_from_bytes<T>, "data"_a, "width"_a, "height"_a);
    c.def_property_readonly("width_blocks", [](const T &t) { return t.bw; });
    c.def_property_readonly("height_blocks", [](const T &t) { return t.bh; });
    c.def_property_readonly("size_blocks", &T::size_blocks);
    c.def("__getitem__", [](const T &t, std::tuple<int, int> k) {
        const int x = py_index(std::get<0>(k), t.bw, "x"), y = py_index(std::get<1>(k), t.bh, "y");
        return t.blocks[(size_t)y * t.bw + x];
    }, "key"_a);
    c.def("__setitem__", [](T &t, std::tuple<int, int> k, const B &v) {
        const int x = py_index(std::get<0>(k), t.bw, "x"), y = py_index(std::get<1>(k), t.bh, "y");
        t.blocks[(size_t)y * t.bw + x] = v;
    }, "key"_a, "value"_a);
}

// ---------------------------------------------------------------------------
// Mip chains and whole-file decodes: the per-level loops of quicktex.dds
// (encode, quicktex/dds.py:282-287; DDSFile.decode_all, dds.py:209-216) as one
// call each.

template <class B>
py::list encode_mips_as(const qtx_job &tmpl, const py::buffer &data, int w, int h, int levels) {
    if (w <= 0) throw std::invalid_argument("Texture width must be greater than 0");
    if (h <= 0) throw std::invalid_argument("Texture height must be greater than 0");
    py::buffer_info info = request_bytes(data, (size_t)w * (size_t)h * 4);
    std::vector<std::unique_ptr<BlockTexture<B>>> outs;
    std::vector<uint8_t *> ptrs;
    std::vector<size_t> caps;
    int lw = w, lh = h;
    for (int m = 0; m < levels; m++) {
        outs.push_back(std::make_unique<BlockTexture<B>>(lw, lh));
        ptrs.push_back(outs.back()->data());
        caps.push_back(outs.back()->nbytes());
        lw = std::max(lw / 2, 1);
        lh = std::max(lh / 2, 1);
    }
    qtx_status st;
    {
        py::gil_scoped_release nogil;
        st = qtx_mips_encode_host(context(), &tmpl, static_cast<const uint8_t *>(info.ptr), w, h, levels, ptrs.data(),
                                  caps.data());
    }
    check(st);
    py::list l;
    for (auto &o : outs) l.append(py::cast(std::move(o)));
    return l;
}

qtx_job encode_template(int format) {
    qtx_job j{};
    j.op = QTX_OP_ENCODE;
    j.format = format;
    return j;
}

// encoder: a BC1/BC3/BC4/BC5 encoder; data: RGBA8 bytes of the base level.
py::list encode_mips(py::object encoder, py::buffer data, int w, int h, int levels) {
    if (py::isinstance<BC1Encoder>(encoder)) {
        qtx_job j = encode_template(QTX_FMT_BC1);
        j.bc1 = encoder.cast<const BC1Encoder &>().p;
        return encode_mips_as<BC1Block>(j, data, w, h, levels);
    }
    if (py::isinstance<BC3Encoder>(encoder)) {
        qtx_job j = encode_template(QTX_FMT_BC3);
        j.bc1 = encoder.cast<const BC3Encoder &>().bc1->p;
        return encode_mips_as<BC3Block>(j, data, w, h, levels);
    }
    if (py::isinstance<BC4Encoder>(encoder)) {
        qtx_job j = encode_template(QTX_FMT_BC4);
        j.chan0 = encoder.cast<const BC4Encoder &>().channel;
        return encode_mips_as<BC4Block>(j, data, w, h, levels);
    }
    if (py::isinstance<BC5Encoder>(encoder)) {
        const BC5Encoder &e = encoder.cast<const BC5Encoder &>();
        qtx_job j = encode_template(QTX_FMT_BC5);
        j.chan0 = e.e0->channel;
        j.chan1 = e.e1->channel;
        return encode_mips_as<BC5Block>(j, data, w, h, levels);
    }
    throw py::type_error("expected a BC1, BC3, BC4 or BC5 encoder");
}

template <class B> py::list decode_all_as(qtx_job tmpl, const py::list &textures) {
    const size_t n = py::len(textures);
    std::vector<std::unique_ptr<RawTexture>> outs;
    std::vector<qtx_job> jobs(n, tmpl);
    std::vector<size_t> caps(n);
    for (size_t i = 0; i < n; i++) {
        const BlockTexture<B> &t = textures[i].cast<const BlockTexture<B> &>();
        outs.push_back(std::make_unique<RawTexture>(t.width, t.height));
        jobs[i].width = t.width;
        jobs[i].height = t.height;
        jobs[i].src = t.cdata();
        jobs[i].dst = outs.back()->data();
        caps[i] = outs.back()->nbytes();
    }
    py::list l;
    if (n == 0) return l;
    qtx_status st;
    {
        py::gil_scoped_release nogil;
        st = qtx_batch_run_host(context(), jobs.data(), caps.data(), n);
    }
    check(st);
    for (auto &o : outs) l.append(py::cast(std::move(o)));
    return l;
}

// `encode auto` (quicktex/cli/encode.py:103-141) for one image: (is_bc3, levels).
py::tuple encode_mips_auto(const BC1Encoder &bc1, py::buffer data, int w, int h, int levels) {
    if (w <= 0) throw std::invalid_argument("Texture width must be greater than 0");
    if (h <= 0) throw std::invalid_argument("Texture height must be greater than 0");
    py::buffer_info info = request_bytes(data, (size_t)w * (size_t)h * 4);
    // staging sized for BC3; the blocks are copied into textures of the chosen format
    std::vector<std::vector<uint8_t>> stage;
    std::vector<uint8_t *> ptrs;
    std::vector<size_t> caps;
    std::vector<std::pair<int, int>> sizes;
    int lw = w, lh = h;
    for (int m = 0; m < levels; m++) {
        sizes.emplace_back(lw, lh);
        stage.emplace_back((size_t)((lw + 3) / 4) * (size_t)((lh + 3) / 4) * 16);
        ptrs.push_back(stage.back().data());
        caps.push_back(stage.back().size());
        lw = std::max(lw / 2, 1);
        lh = std::max(lh / 2, 1);
    }
    int32_t fmt = 0;
    qtx_status st;
    {
        py::gil_scoped_release nogil;
        st = qtx_mips_encode_auto_host(context(), &bc1.p, static_cast<const uint8_t *>(info.ptr), w, h, levels,
                                       ptrs.data(), caps.data(), &fmt);
    }
    check(st);
    py::list l;
    for (int m = 0; m < levels; m++) {
        if (fmt == QTX_FMT_BC3) {
            auto t = std::make_unique<BlockTexture<BC3Block>>(sizes[m].first, sizes[m].second);
            std::memcpy(t->data(), stage[m].data(), t->nbytes());
            l.append(py::cast(std::move(t)));
        } else {
            auto t = std::make_unique<BlockTexture<BC1Block>>(sizes[m].first, sizes[m].second);
            std::memcpy(t->data(), stage[m].data(), t->nbytes());
            l.append(py::cast(std::move(t)));
        }
    }
    return py::make_tuple(fmt == QTX_FMT_BC3, l);
}

py::bytes swizzle_bc3nm(py::buffer data, int w, int h) {
    if (w <= 0 || h <= 0) throw std::invalid_argument("Texture size must be greater than 0");
    const size_t n = (size_t)w * (size_t)h * 4;
    py::buffer_info info = request_bytes(data, n);
    std::string out(n, '\0');
    qtx_status st;
    {
        py::gil_scoped_release nogil;
        st = qtx_swizzle_bc3nm_host(context(), static_cast<const uint8_t *>(info.ptr), w, h,
                                    reinterpret_cast<uint8_t *>(&out[0]));
    }
    check(st);
    return py::bytes(out);
}

qtx_job decode_template(int format) {
    qtx_job j{};
    j.op = QTX_OP_DECODE;
    j.format = format;
    return j;
}

// decoder: a BC1/BC3/BC4/BC5 decoder; textures: block textures of its format.
py::list decode_all(py::object decoder, py::list textures) {
    if (py::isinstance<BC1Decoder>(decoder)) {
        const BC1Decoder &d = decoder.cast<const BC1Decoder &>();
        qtx_job j = decode_template(QTX_FMT_BC1);
        j.bc1.interpolator = d.interpolator->type();
        j.write_alpha = d.write_alpha;
        return decode_all_as<BC1Block>(j, textures);
    }
    if (py::isinstance<BC3Decoder>(decoder)) {
        qtx_job j = decode_template(QTX_FMT_BC3);
        j.bc1.interpolator = decoder.cast<const BC3Decoder &>().bc1->interpolator->type();
        return decode_all_as<BC3Block>(j, textures);
    }
    if (py::isinstance<BC4Decoder>(decoder)) {
        qtx_job j = decode_template(QTX_FMT_BC4);
        j.chan0 = decoder.cast<const BC4Decoder &>().channel;
        return decode_all_as<BC4Block>(j, textures);
    }
    if (py::isinstance<BC5Decoder>(decoder)) {
        const BC5Decoder &d = decoder.cast<const BC5Decoder &>();
        qtx_job j = decode_template(QTX_FMT_BC5);
        j.chan0 = d.d0->channel;
        j.chan1 = d.d1->channel;
        return decode_all_as<BC5Block>(j, textures);
    }
    throw py::type_error("expected a BC1, BC3, BC4 or BC5 decoder");
}

}  // namespace qb

PYBIND11_MODULE(_quicktex, m) {
    using namespace qb;
    m.doc() = "MI355X (gfx950) S3TC codec: quicktex's Python API over HIP kernels";
    m.attr("__version__") = "0.3.1+mi355x";

    py::class_<Texture> texture(m, "Texture", py::buffer_protocol());
    texture.def_property_readonly("nbytes", &Texture::nbytes);
    texture.def_property_readonly("size", &Texture::size);
    texture.def_property_readonly("width", [](const Texture &t) { return t.width; });
    texture.def_property_readonly("height", [](const Texture &t) { return t.height; });
    texture.def_buffer([](Texture &t) { return py::buffer_info(t.data(), (py::ssize_t)t.nbytes()); });
    texture.def("tobytes", [](const Texture &t) { return py::bytes(reinterpret_cast<const char *>(t.cdata()), t.nbytes()); });

    py::class_<RawTexture, Texture> raw(m, "RawTexture");
    raw.def(py::init<int, int>(), "width"_a, "height"_a);
    raw.def_static("frombytes", &texture_from_bytes<RawTexture>, "data"_a, "width"_a, "height"_a);
    raw.def("__getitem__", [](const RawTexture &t, std::tuple<int, int> k) {
        return t.get(py_index(std::get<0>(k), t.width, "x"), py_index(std::get<1>(k), t.height, "y"));
    }, "key"_a);
    raw.def("__setitem__", [](RawTexture &t, std::tuple<int, int> k, const Color &c) {
        t.set(py_index(std::get<0>(k), t.width, "x"), py_index(std::get<1>(k), t.height, "y"), c);
    }, "key"_a, "value"_a);

    py::module_ s3tc = m.def_submodule("_s3tc", "s3tc codecs on gfx950");

    // interpolators (quicktex/s3tc/interpolator/_bindings.cpp:24-54), constructible here
    py::module_ mi = s3tc.def_submodule("_interpolator", "interpolator tags");
    py::class_<Interpolator, std::shared_ptr<Interpolator>> ideal(mi, "Interpolator");
    ideal.def(py::init<>());
    py::class_<InterpolatorRound, Interpolator, std::shared_ptr<InterpolatorRound>>(mi, "InterpolatorRound").def(py::init<>());
    py::class_<InterpolatorNvidia, Interpolator, std::shared_ptr<InterpolatorNvidia>>(mi, "InterpolatorNvidia").def(py::init<>());
    py::class_<InterpolatorAMD, Interpolator, std::shared_ptr<InterpolatorAMD>>(mi, "InterpolatorAMD").def(py::init<>());

    // BC1 (quicktex/s3tc/bc1/_bindings.cpp:44-190)
    py::module_ m1 = s3tc.def_submodule("_bc1", "BC1");
    auto b1 = bind_block<BC1Block>(m1, "BC1Block");
    b1.def(py::init<>());
    b1.def(py::init<Color, Color, Sel4x4>(), "color0"_a, "color1"_a, "selectors"_a);
    b1.def_property("endpoints", &BC1Block::endpoints, &BC1Block::set_endpoints);
    b1.def_property("selectors", &BC1Block::selectors, &BC1Block::set_selectors);
    b1.def_property_readonly("is_3color", &BC1Block::is_3color);
    bind_block_texture<BC1Block>(m1, "BC1Texture");

    py::class_<BC1Encoder, std::shared_ptr<BC1Encoder>> e1(m1, "BC1Encoder");
    py::enum_<EndpointMode>(e1, "EndpointMode")
        .value("LeastSquares", EndpointMode::LeastSquares)
        .value("BoundingBox", EndpointMode::BoundingBox)
        .value("BoundingBoxInt", EndpointMode::BoundingBoxInt)
        .value("PCA", EndpointMode::PCA);
    py::enum_<ErrorMode>(e1, "ErrorMode")
        .value("None", ErrorMode::None)
        .value("Faster", ErrorMode::Faster)
        .value("Check2", ErrorMode::Check2)
        .value("Full", ErrorMode::Full);
    py::enum_<ColorMode>(e1, "ColorMode")
        .value("FourColor", ColorMode::FourColor)
        .value("ThreeColor", ColorMode::ThreeColor)
        .value("ThreeColorBlack", ColorMode::ThreeColorBlack);
    e1.def(py::init([](unsigned level, ColorMode mode) { return std::make_shared<BC1Encoder>(level, mode, nullptr); }),
           "level"_a = 5, "color_mode"_a = ColorMode::FourColor);
    e1.def(py::init<unsigned, ColorMode, InterpPtr>(), "level"_a, "color_mode"_a, "interpolator"_a);
    e1.def("encode", &BC1Encoder::encode, "texture"_a);
    e1.def("set_level", &BC1Encoder::set_level, "level"_a);
    e1.def_property_readonly("interpolator", [](const BC1Encoder &e) { return e.interpolator; });
    e1.def_property_readonly("color_mode", [](const BC1Encoder &e) { return e.color_mode; });
    e1.def_property("error_mode", [](const BC1Encoder &e) { return (ErrorMode)e.p.error_mode; },
                    [](BC1Encoder &e, ErrorMode v) { e.p.error_mode = (int32_t)v; });
    e1.def_property("endpoint_mode", [](const BC1Encoder &e) { return (EndpointMode)e.p.endpoint_mode; },
                    [](BC1Encoder &e, EndpointMode v) { e.p.endpoint_mode = (int32_t)v; });
    e1.def_property("two_ls_passes", [](const BC1Encoder &e) { return e.p.two_ls_passes != 0; },
                    [](BC1Encoder &e, bool v) { e.p.two_ls_passes = v; });
    e1.def_property("two_ep_passes", [](const BC1Encoder &e) { return e.p.two_ep_passes != 0; },
                    [](BC1Encoder &e, bool v) { e.p.two_ep_passes = v; });
    e1.def_property("two_cf_passes", [](const BC1Encoder &e) { return e.p.two_cf_passes != 0; },
                    [](BC1Encoder &e, bool v) { e.p.two_cf_passes = v; });
    e1.def_property("exhaustive", [](const BC1Encoder &e) { return e.p.exhaustive != 0; },
                    [](BC1Encoder &e, bool v) { e.p.exhaustive = v; });
    e1.def_property("search_rounds", [](const BC1Encoder &e) { return e.p.search_rounds; },
                    [](BC1Encoder &e, unsigned v) { e.p.search_rounds = v; });
    e1.def_property("orderings", &BC1Encoder::orderings, &BC1Encoder::set_orderings);
    e1.def_property_readonly_static("max_power_iterations", [](py::object) { return 10u; });
    e1.def_property_readonly_static("min_power_iterations", [](py::object) { return 4u; });
    e1.def_property("power_iterations", [](const BC1Encoder &e) { return e.p.power_iterations; },
                    &BC1Encoder::set_power_iterations);

    py::class_<BC1Decoder, std::shared_ptr<BC1Decoder>> d1(m1, "BC1Decoder");
    d1.def(py::init([](bool wa) { return std::make_shared<BC1Decoder>(wa, nullptr); }), "write_alpha"_a = false);
    d1.def(py::init<bool, InterpPtr>(), "write_alpha"_a, "interpolator"_a);
    d1.def("decode", &BC1Decoder::decode, "texture"_a);
    d1.def_property_readonly("interpolator", [](const BC1Decoder &d) { return d.interpolator; });
    d1.def_readwrite("write_alpha", &BC1Decoder::write_alpha);

    // BC4 (quicktex/s3tc/bc4/_bindings.cpp:40-121)
    py::module_ m4 = s3tc.def_submodule("_bc4", "BC4");
    auto b4 = bind_block<BC4Block>(m4, "BC4Block");
    b4.def(py::init<>());
    b4.def(py::init<uint8_t, uint8_t, Sel4x4>(), "endpoint0"_a, "endpoint1"_a, "selectors"_a);
    b4.def_property("endpoints", &BC4Block::endpoints, &BC4Block::set_endpoints);
    b4.def_property("selectors", &BC4Block::selectors, &BC4Block::set_selectors);
    b4.def_property_readonly("values", &BC4Block::values);
    b4.def_property_readonly("is_6value", &BC4Block::is_6value);
    bind_block_texture<BC4Block>(m4, "BC4Texture");
    py::class_<BC4Encoder, std::shared_ptr<BC4Encoder>> e4(m4, "BC4Encoder");
    e4.def(py::init<uint8_t>(), "channel"_a = 3);
    e4.def("encode", &BC4Encoder::encode, "texture"_a);
    e4.def_property_readonly("channel", [](const BC4Encoder &e) { return e.channel; });
    py::class_<BC4Decoder, std::shared_ptr<BC4Decoder>> d4(m4, "BC4Decoder");
    d4.def(py::init<uint8_t>(), "channel"_a = 3);
    d4.def("decode", &BC4Decoder::decode, "texture"_a);
    d4.def_property_readonly("channel", [](const BC4Decoder &d) { return d.channel; });

    // BC3 (quicktex/s3tc/bc3/_bindings.cpp:44-119)
    py::module_ m3 = s3tc.def_submodule("_bc3", "BC3");
    auto b3 = bind_block<BC3Block>(m3, "BC3Block");
    b3.def(py::init<>());
    b3.def(py::init<BC4Block, BC1Block>(), "alpha_block"_a, "color_block"_a);
    b3.def_readwrite("alpha_block", &BC3Block::alpha_block);
    b3.def_readwrite("color_block", &BC3Block::color_block);
    b3.def_property("blocks", &BC3Block::blocks, &BC3Block::set_blocks);
    bind_block_texture<BC3Block>(m3, "BC3Texture");
    py::class_<BC3Encoder, std::shared_ptr<BC3Encoder>> e3(m3, "BC3Encoder");
    e3.def(py::init([](unsigned level) { return std::make_shared<BC3Encoder>(level, nullptr); }), "level"_a = 5);
    e3.def(py::init<unsigned, InterpPtr>(), "level"_a, "interpolator"_a);
    e3.def("encode", &BC3Encoder::encode, "texture"_a);
    e3.def_property_readonly("bc1_encoder", [](const BC3Encoder &e) { return e.bc1; });
    e3.def_property_readonly("bc4_encoder", [](const BC3Encoder &e) { return e.bc4; });
    py::class_<BC3Decoder, std::shared_ptr<BC3Decoder>> d3(m3, "BC3Decoder");
    d3.def(py::init([]() { return std::make_shared<BC3Decoder>(nullptr); }));
    d3.def(py::init<InterpPtr>(), "interpolator"_a);
    d3.def("decode", &BC3Decoder::decode, "texture"_a);
    d3.def_property_readonly("bc1_decoder", [](const BC3Decoder &d) { return d.bc1; });
    d3.def_property_readonly("bc4_decoder", [](const BC3Decoder &d) { return d.bc4; });

    // BC5 (quicktex/s3tc/bc5/_bindings.cpp:37-108)
    py::module_ m5 = s3tc.def_submodule("_bc5", "BC5");
    auto b5 = bind_block<BC5Block>(m5, "BC5Block");
    b5.def(py::init<>());
    b5.def(py::init<BC4Block, BC4Block>(), "chan0_block"_a, "chan1_block"_a);
    b5.def_readwrite("chan0_block", &BC5Block::chan0_block);
    b5.def_readwrite("chan1_block", &BC5Block::chan1_block);
    b5.def_property("blocks", &BC5Block::blocks, &BC5Block::set_blocks);
    bind_block_texture<BC5Block>(m5, "BC5Texture");
    py::class_<BC5Encoder, std::shared_ptr<BC5Encoder>> e5(m5, "BC5Encoder");
    e5.def(py::init<uint8_t, uint8_t>(), "chan0"_a = 0, "chan1"_a = 1);
    e5.def("encode", &BC5Encoder::encode, "texture"_a);
    e5.def_property_readonly("channels", [](const BC5Encoder &e) { return std::make_tuple(e.e0->channel, e.e1->channel); });
    e5.def_property_readonly("bc4_encoders", [](const BC5Encoder &e) { return std::make_tuple(e.e0, e.e1); });
    py::class_<BC5Decoder, std::shared_ptr<BC5Decoder>> d5(m5, "BC5Decoder");
    d5.def(py::init<uint8_t, uint8_t>(), "chan0"_a = 0, "chan1"_a = 1);
    d5.def("decode", &BC5Decoder::decode, "texture"_a);
    d5.def_property_readonly("channels", [](const BC5Decoder &d) { return std::make_tuple(d.d0->channel, d.d1->channel); });
    d5.def_property_readonly("bc4_decoders", [](const BC5Decoder &d) { return std::make_tuple(d.d0, d.d1); });

    // whole mip chains (quicktex.dds)
    m.def("_mip_count", [](int w, int h, int requested) { return qtx_mip_count(w, h, requested); }, "width"_a,
          "height"_a, "requested"_a = 0);
    m.def("_encode_mips", &encode_mips, "encoder"_a, "data"_a, "width"_a, "height"_a, "levels"_a);
    m.def("_decode_all", &decode_all, "decoder"_a, "textures"_a);
    m.def("_encode_mips_auto", &encode_mips_auto, "bc1_encoder"_a, "data"_a, "width"_a, "height"_a, "levels"_a);
    m.def("_swizzle_bc3nm", &swizzle_bc3nm, "data"_a, "width"_a, "height"_a);
}
