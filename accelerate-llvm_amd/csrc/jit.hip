// Run-time compiled skeletons: a skeleton node whose scalar functions come
// from the program (text, see jit_expr.hpp) is translated to HIP C++ around
// the templates of jit/skel.hpp, compiled for gfx950 with hiprtc, cached on
// disk by a hash of its source, loaded as a module and launched.
//
// This is the reference's compile/link/execute of one skeleton node:
//   compile  PTX/Compile.hs:84-165 (llvmOfPreOpenAcc, then clang to a cubin),
//   cache    LLVM/Compile/Cache.hs:57-91 (keyed by the node's hash, stored
//            under the user's cache directory),
//   link     PTX/Link.hs:57-121 (load the module, look the kernels up),
//   embed    PTX/Embed.hs:50-79 (a compiled image carried by the program and
//            loaded without compiling: acc_hip_kernel_image/_from_image),
//   execute  PTX/Execute.hs:222-776 (dispatch per skeleton: output sizes,
//            empty cases, clone versus in place, launch configuration).
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>
#include <chrono>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>
#include "host.hpp"
#include "jit_expr.hpp"
#include "jit/args.hpp"

namespace acc {
namespace jit {

// The device headers the generated source includes (their text, embedded at
// build time).
struct HeaderText {
  const char* name;
  const char* text;
};
static const HeaderText kHeaders[] = {
#include "jit_headers.inc"
};
constexpr int kNumHeaders = (int)(sizeof(kHeaders) / sizeof(kHeaders[0]));

enum Kind { K_FOLD, K_SCAN, K_FOLDSEG, K_PERMUTE, K_STENCIL, K_GENERATE };
static const char* kKindName[] = {"fold", "scan", "foldseg", "permute", "stencil", "generate"};
enum Bound { B_CLAMP, B_MIRROR, B_WRAP, B_CONSTANT, B_FUNCTION };

struct Node {
  Kind kind = K_FOLD;
  std::vector<int> elt;          // element fields (stencil: the output)
  int stencil_in = F32;          // stencil input element
  std::vector<int> free;         // free array element types
  bool has_seed = false, delayed = false, prime = false, rev = false;
  int off_type = I64;            // foldSeg offsets
  int ry = 1, rx = 1;            // stencil radii
  int bound = B_CLAMP;
  int K = 4;                     // scan: vectors per lane
  bool cas = false;              // permute: compare-and-swap update (else a lock per slot)
  int gen_arity = 1;
  Layout lay;
  std::string source;
  std::string key;               // 32 hex digits
};

static const char* kOptions[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                                 "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
constexpr int kNumOptions = (int)(sizeof(kOptions) / sizeof(kOptions[0]));
static const char* kVersion = "accelerate-hip-jit-1";

// --------------------------------------------------------- node -> source
static std::vector<int> parse_types(const SX& item, size_t from, const char* what) {
  std::vector<int> r;
  for (size_t k = from; k < item.l.size(); ++k) {
    if (!item.l[k].atom) throw Error(std::string(what) + ": expected type names");
    const int t = sc_parse(item.l[k].a);
    if (t < 0) throw Error(std::string(what) + ": unknown type " + item.l[k].a);
    r.push_back(t);
  }
  return r;
}

static std::vector<std::string> fields_of(const std::string& v, size_t n) {
  std::vector<std::string> r;
  for (size_t f = 0; f < n; ++f) r.push_back(v + ".f" + std::to_string(f));
  return r;
}

static std::string ptr_list(const char* arr, int n = accjit::kMaxFields) {
  std::string s = "{";
  for (int k = 0; k < n; ++k) s += std::string(k ? ", " : "") + arr + "[" + std::to_string(k) + "]";
  return s + "}";
}

// manifest loader / storer of struct `E` with fields `f`
static std::string manifest_loader(const char* name, const char* E, const std::vector<int>& f) {
  std::string s = std::string("struct ") + name + " {\n  const void* p[4];\n  uint32_t vec;\n";
  s += std::string("  __device__ __forceinline__ ") + E + " operator()(int64_t i) const {\n    " + E + " e;\n";
  for (size_t k = 0; k < f.size(); ++k)
    s += "    e.f" + std::to_string(k) + " = ((const " + sc_storetype(f[k]) + "*)p[" + std::to_string(k) + "])[i];\n";
  s += "    return e;\n  }\n";
  s += std::string("  __device__ __forceinline__ void four(int64_t i, ") + E + "* o) const {\n    if (vec) {\n";
  for (size_t k = 0; k < f.size(); ++k) {
    const std::string T = sc_storetype(f[k]), ks = std::to_string(k);
    s += "      { struct alignas(" + std::to_string(4 * sc_size(f[k])) + ") V { " + T + " e[4]; }; const V v = *(const V*)((const " +
         T + "*)p[" + ks + "] + i);\n        for (int j = 0; j < 4; ++j) o[j].f" + ks + " = v.e[j]; }\n";
  }
  s += "    } else {\n      for (int j = 0; j < 4; ++j) o[j] = (*this)(i + j);\n    }\n  }\n};\n";
  return s;
}

static std::string manifest_storer(const char* name, const char* E, const std::vector<int>& f) {
  std::string s = std::string("struct ") + name + " {\n  void* p[4];\n  uint32_t vec;\n";
  s += std::string("  __device__ __forceinline__ void operator()(int64_t i, const ") + E + "& e) const {\n";
  for (size_t k = 0; k < f.size(); ++k)
    s += "    ((" + std::string(sc_storetype(f[k])) + "*)p[" + std::to_string(k) + "])[i] = e.f" + std::to_string(k) + ";\n";
  s += "  }\n";
  s += std::string("  __device__ __forceinline__ void four(int64_t i, const ") + E + "* y) const {\n    if (vec) {\n";
  for (size_t k = 0; k < f.size(); ++k) {
    const std::string T = sc_storetype(f[k]), ks = std::to_string(k);
    s += "      { struct alignas(" + std::to_string(4 * sc_size(f[k])) + ") V { " + T + " e[4]; }; V v;\n        for (int j = 0; j < 4; ++j) v.e[j] = y[j].f" +
         ks + ";\n        *(V*)((" + T + "*)p[" + ks + "] + i) = v; }\n";
  }
  s += "    } else {\n      for (int j = 0; j < 4; ++j) (*this)(i + j, y[j]);\n    }\n  }\n};\n";
  return s;
}

static const SX* find_item(const SX& node, const char* head) {
  for (size_t k = 1; k < node.l.size(); ++k)
    if (node.l[k].is(head)) return &node.l[k];
  return nullptr;
}

static bool has_flag(const SX& node, const char* flag) {
  for (size_t k = 1; k < node.l.size(); ++k) {
    const SX& it = node.l[k];
    if ((it.atom && it.a == flag) || (!it.atom && it.l.size() == 1 && it.l[0].atom && it.l[0].a == flag)) return true;
  }
  return false;
}

static const SX& item_fun(const SX* item, const char* what) {
  if (!item || item->l.size() != 2) throw Error(std::string(what) + ": expected (" + what + " (fun ...))");
  return item->l[1];
}

static uint64_t fnv1a(const std::string& s, uint64_t h) {
  for (unsigned char c : s) {
    h ^= c;
    h *= 0x100000001b3ull;
  }
  return h;
}

static std::string node_key(const std::string& source) {
  std::string all = std::string(kVersion) + "\n";
  for (int i = 0; i < kNumOptions; ++i) all += std::string(kOptions[i]) + "\n";
  for (int i = 0; i < kNumHeaders; ++i) all += std::string(kHeaders[i].name) + "\n" + kHeaders[i].text;
  all += source;
  char b[40];
  snprintf(b, sizeof b, "%016llx%016llx", (unsigned long long)fnv1a(all, 0xcbf29ce484222325ull),
           (unsigned long long)fnv1a(all, 0x84222325cbf29ce4ull));
  return b;
}

static void build_node(const std::string& text, Node& nd) {
  const SX node = parse_sx(text);
  if (node.atom || node.l.empty() || !node.l[0].atom) throw Error("expected (skeleton items...)");
  const std::string& kind = node.l[0].a;
  int ki = -1;
  for (int k = 0; k < 6; ++k)
    if (kind == kKindName[k]) ki = k;
  if (ki < 0) throw Error("unknown skeleton " + kind + " (fold, scan, foldseg, permute, stencil, generate)");
  nd.kind = (Kind)ki;
  if (const SX* fr = find_item(node, "free")) nd.free = parse_types(*fr, 1, "free");
  if ((int)nd.free.size() > accjit::kMaxFree) throw Error("at most 8 free arrays");
  Emitter proto;
  proto.free_types = nd.free;
  proto.foreign = std::make_shared<std::map<std::string, std::string>>();

  std::string src = "#include \"jit/prelude.hpp\"\n#include \"jit/skel.hpp\"\nusing namespace accjit;\n";
  auto fn_struct = [&](const char* name, const std::string& sig, const std::string& body) {
    return std::string("struct ") + name + " {\n  Free fv;\n  __device__ __forceinline__ " + sig + " const {\n" + body +
           "  }\n};\n";
  };

  if (nd.kind == K_STENCIL) {
    const SX* in = find_item(node, "in");
    const SX* out = find_item(node, "out");
    const SX* rad = find_item(node, "radius");
    if (!in || !out || !rad) throw Error("stencil: needs (in T) (out T) (radius RY RX)");
    const std::vector<int> ti = parse_types(*in, 1, "in"), to = parse_types(*out, 1, "out");
    if (ti.size() != 1 || to.size() != 1) throw Error("stencil: scalar input and output elements");
    nd.stencil_in = ti[0];
    nd.elt = to;
    if (rad->l.size() != 3 || !rad->l[1].atom || !rad->l[2].atom) throw Error("stencil: (radius RY RX)");
    nd.ry = atoi(rad->l[1].a.c_str());
    nd.rx = atoi(rad->l[2].a.c_str());
    if (nd.ry < 0 || nd.ry > 4 || nd.rx < 0 || nd.rx > 4) throw Error("stencil: radii 0..4");
    const std::string Ti = sc_storetype(ti[0]), To = sc_storetype(to[0]);
    Emitter em = proto;
    em.stencil = true;
    em.nb_ry = nd.ry;
    em.nb_rx = nd.rx;
    em.nb_stride = 64 + 2 * nd.rx;
    em.nb_type = ti[0];
    const SX* fun = nullptr;
    for (size_t k = 1; k < node.l.size(); ++k)
      if (node.l[k].is("fun")) fun = &node.l[k];
    if (!fun) throw Error("stencil: needs the stencil function (fun () E)");
    src += fn_struct("Fs", To + " operator()(const " + Ti + "* nb)", emit_function(*fun, {}, {}, to, "", em));
    // boundary
    const SX* b = find_item(node, "boundary");
    std::string bbody;
    if (!b || b->l.size() != 2) throw Error("stencil: (boundary clamp|mirror|wrap|(constant E)|(function (fun (y x) E)))");
    const SX& bv = b->l[1];
    if (bv.atom && bv.a == "clamp") nd.bound = B_CLAMP, bbody = "    return in[map_clamp(y, H) * W + map_clamp(x, W)];\n";
    else if (bv.atom && bv.a == "mirror") nd.bound = B_MIRROR, bbody = "    return in[map_mirror(y, H) * W + map_mirror(x, W)];\n";
    else if (bv.atom && bv.a == "wrap") nd.bound = B_WRAP, bbody = "    return in[map_wrap(y, H) * W + map_wrap(x, W)];\n";
    else if (bv.is("constant") && bv.l.size() == 2) {
      nd.bound = B_CONSTANT;
      SX f;
      f.l.resize(3);
      f.l[0].atom = true;
      f.l[0].a = "fun";
      f.l[2] = bv.l[1];
      bbody = emit_function(f, {}, {}, ti, "", proto);
    } else if (bv.is("function") && bv.l.size() == 2) {
      nd.bound = B_FUNCTION;
      bbody = emit_function(bv.l[1], {{I64}, {I64}}, {{"y"}, {"x"}}, ti, "", proto);
    } else {
      throw Error("stencil: unknown boundary " + bv.str());
    }
    src += fn_struct("Bnd", Ti + " operator()(const " + Ti + "* in, int64_t y, int64_t x, int64_t H, int64_t W)", bbody);
    src += "extern \"C\" __global__ __launch_bounds__(256) void acc_jit_stencil(Args a) {\n  stencil2d<" + Ti + ", " + To +
           ", Fs, Bnd, " + std::to_string(nd.ry) + ", " + std::to_string(nd.rx) + ">(a, Fs{a.fv}, Bnd{a.fv});\n}\n";
    // foreign scalar functions, at file scope ahead of the functors that call them
  std::string defs;
  for (const auto& kv : *proto.foreign) defs += kv.second;
  const std::string head = "using namespace accjit;\n";
  const size_t at = src.find(head) + head.size();
  nd.source = src.substr(0, at) + defs + src.substr(at);
    nd.key = node_key(src);
    return;
  }

  const SX* el = find_item(node, "elt");
  if (!el) throw Error(kind + ": needs (elt T...)");
  nd.elt = parse_types(*el, 1, "elt");
  if (nd.elt.empty() || (int)nd.elt.size() > accjit::kMaxFields) throw Error("elt: 1..4 fields");
  nd.lay = layout_of(nd.elt);
  src += struct_decl("E", nd.elt);
  src += "static_assert(sizeof(E) == " + std::to_string(nd.lay.size) + " && sizeof(acc::Maybe<E>) == " +
         std::to_string(nd.lay.maybe_size) + ", \"element layout\");\n";
  const auto ef = fields_of("a", nd.elt.size()), bf = fields_of("b", nd.elt.size());

  if (nd.kind == K_GENERATE) {
    const SX* fun = nullptr;
    for (size_t k = 1; k < node.l.size(); ++k)
      if (node.l[k].is("fun")) fun = &node.l[k];
    if (!fun || fun->l.size() != 3 || fun->l[1].atom) throw Error("generate: needs (fun (i) E) or (fun (r c) E)");
    nd.gen_arity = (int)fun->l[1].l.size();
    std::string body;
    if (nd.gen_arity == 1) body = emit_function(*fun, {{I64}}, {{"i"}}, nd.elt, "E", proto);
    else if (nd.gen_arity == 2)
      body = "    const int64_t row_ = i / inner, col_ = i - row_ * inner;\n" +
             emit_function(*fun, {{I64}, {I64}}, {{"row_"}, {"col_"}}, nd.elt, "E", proto);
    else throw Error("generate: a function of the linear index (i) or of (row column)");
    src += fn_struct("Gen", "E operator()(int64_t i, int64_t inner)", body);
    src += manifest_storer("St", "E", nd.elt);
    src += "extern \"C\" __global__ __launch_bounds__(256) void acc_jit_generate(Args a) {\n  generate<E>(a, Gen{a.fv}, St{" +
           ptr_list("a.out") + ", a.vec});\n}\n";
    // foreign scalar functions, at file scope ahead of the functors that call them
  std::string defs;
  for (const auto& kv : *proto.foreign) defs += kv.second;
  const std::string head = "using namespace accjit;\n";
  const size_t at = src.find(head) + head.size();
  nd.source = src.substr(0, at) + defs + src.substr(at);
    nd.key = node_key(src);
    return;
  }

  // combine function (permute: new, old)
  const SX* fun = nullptr;
  for (size_t k = 1; k < node.l.size(); ++k)
    if (node.l[k].is("fun")) fun = &node.l[k];
  if (!fun) throw Error(kind + ": needs the combine function (fun (a b) E)");
  src += fn_struct("Fn", "E operator()(const E& a, const E& b)", emit_function(*fun, {nd.elt, nd.elt}, {ef, bf}, nd.elt, "E", proto));
  // seed
  const SX* sd = find_item(node, "seed");
  nd.has_seed = sd != nullptr;
  if (sd) {
    if (sd->l.size() != 2) throw Error("seed: expected (seed E)");
    SX f;
    f.l.resize(3);
    f.l[0].atom = true;
    f.l[0].a = "fun";
    f.l[2] = sd->l[1];
    src += "struct Seed {\n  Free fv;\n  static constexpr bool has = true;\n  __device__ __forceinline__ E operator()() const {\n" +
           emit_function(f, {}, {}, nd.elt, "E", proto) + "  }\n};\n";
  } else {
    src += "struct Seed {\n  Free fv;\n  static constexpr bool has = false;\n  __device__ __forceinline__ E operator()() const { return E{}; }\n};\n";
  }
  // input: manifest fields, or a delayed producer
  const SX* inp = find_item(node, "input");
  nd.delayed = inp != nullptr;
  if (inp) {
    const std::string body = emit_function(item_fun(inp, "input"), {{I64}}, {{"i"}}, nd.elt, "E", proto);
    src += "struct Ld {\n  Free fv;\n  __device__ __forceinline__ E operator()(int64_t i) const {\n" + body +
           "  }\n  __device__ __forceinline__ void four(int64_t i, E* o) const {\n    for (int j = 0; j < 4; ++j) o[j] = (*this)(i + j);\n  }\n};\n";
  } else {
    src += manifest_loader("Ld", "E", nd.elt);
  }
  const std::string LD = nd.delayed ? "Ld{a.fv}" : "Ld{" + ptr_list("a.in") + ", a.vec}";
  const std::string ST = "St{" + ptr_list("a.out") + ", a.vec}";
  const std::string FN = "Fn{a.fv}", SEED = "Seed{a.fv}";

  if (nd.kind == K_FOLD || nd.kind == K_FOLDSEG) {
    src += manifest_storer("St", "E", nd.elt);
    if (nd.kind == K_FOLD) {
      src += "extern \"C\" __global__ __launch_bounds__(256) void acc_jit_fold_all(Args a) {\n  fold_all<E>(a, " + FN + ", " +
             SEED + ", " + LD + ", " + ST + ");\n}\n";
      for (int wv = 0; wv < 2; ++wv)
        src += std::string("extern \"C\" __global__ __launch_bounds__(256) void acc_jit_fold_seg_") + (wv ? "wave" : "block") +
               "(Args a) {\n  fold_seg<E, Fn, Seed, Ld, St, OffUniform, " + (wv ? "true" : "false") + ">(a, " + FN + ", " +
               SEED + ", " + LD + ", " + ST + ", OffUniform{a.inner});\n}\n";
    } else {
      const SX* of = find_item(node, "offsets");
      if (of) {
        const std::vector<int> ot = parse_types(*of, 1, "offsets");
        if (ot.size() != 1 || (ot[0] != I32 && ot[0] != I64)) throw Error("offsets: i32 or i64");
        nd.off_type = ot[0];
      }
      const std::string OT = nd.off_type == I32 ? "int32_t" : "int64_t";
      for (int wv = 0; wv < 2; ++wv)
        src += std::string("extern \"C\" __global__ __launch_bounds__(256) void acc_jit_fold_seg_") + (wv ? "wave" : "block") +
               "(Args a) {\n  fold_seg<E, Fn, Seed, Ld, St, OffArray<" + OT + ">, " + (wv ? "true" : "false") + ">(a, " + FN +
               ", " + SEED + ", " + LD + ", " + ST + ", OffArray<" + OT + ">{(const " + OT + "*)a.offsets});\n}\n";
    }
  } else if (nd.kind == K_SCAN) {
    const SX* d = find_item(node, "dir");
    if (d) {
      if (d->l.size() != 2 || !d->l[1].atom || (d->l[1].a != "l2r" && d->l[1].a != "r2l"))
        throw Error("scan: (dir l2r|r2l)");
      nd.rev = d->l[1].a == "r2l";
    }
    nd.prime = has_flag(node, "prime");
    if (nd.prime && !nd.has_seed) throw Error("scan': needs a seed");
    const int words = (nd.lay.size + 3) / 4;
    nd.K = words <= 1 ? 8 : (words <= 2 ? 4 : (words <= 4 ? 2 : 1));
    src += manifest_storer("St", "E", nd.elt);
    // totals / the seed at the end of a row
    src += "struct Tot {\n  void* p[4];\n  int64_t at, row;\n  __device__ __forceinline__ void operator()(const E& e) const { put(at, e); }\n"
           "  __device__ __forceinline__ void operator()(int64_t r, const E& e) const { put(r * row + at, e); }\n"
           "  __device__ __forceinline__ void put(int64_t i, const E& e) const {\n";
    for (size_t k = 0; k < nd.elt.size(); ++k)
      src += "    ((" + std::string(sc_storetype(nd.elt[k])) + "*)p[" + std::to_string(k) + "])[i] = e.f" + std::to_string(k) + ";\n";
    src += "  }\n};\n";
    const std::string R = nd.rev ? "true" : "false";
    // Tot's target: sums (scan') or the output row (scanl z / scanr z); `at`
    // and `row` travel in chunk / out_row
    const std::string TOT = nd.prime ? "Tot{" + ptr_list("a.sums") + ", 0, 1}" : "Tot{" + ptr_list("a.out") + ", a.chunk, a.out_row}";
    src += "extern \"C\" __global__ __launch_bounds__(256) void acc_jit_scan(Args a) {\n  scan_all<E, Fn, Seed, Ld, St, Tot, " +
           std::to_string(nd.K) + ", " + R + ">(a, " + FN + ", " + SEED + ", " + LD + ", " + ST + ", " + TOT +
           ", a.exclusive != 0);\n}\n";
    src += "extern \"C\" __global__ __launch_bounds__(256) void acc_jit_scan_rows(Args a) {\n  scan_rows<E, Fn, Seed, Ld, St, Tot, " +
           R + ">(a, " + FN + ", " + SEED + ", " + LD + ", " + ST + ", " + TOT + ", a.exclusive != 0, a.out_row);\n}\n";
  } else if (nd.kind == K_PERMUTE) {
    const SX* pj = find_item(node, "project");
    const std::string pbody = emit_function(item_fun(pj, "project"), {{I64}}, {{"i"}}, {I64}, "", proto);
    src += "struct Proj {\n  Free fv;\n  __device__ __forceinline__ int64_t operator()(int64_t i) const {\n" + pbody + "  }\n};\n";
    nd.cas = nd.elt.size() == 1 && (sc_size(nd.elt[0]) == 4 || sc_size(nd.elt[0]) == 8);
    src += "struct Out {\n  void* p[4];\n  __device__ __forceinline__ E load(int64_t j) const {\n    E e;\n";
    for (size_t k = 0; k < nd.elt.size(); ++k)
      src += "    e.f" + std::to_string(k) + " = ((const " + sc_storetype(nd.elt[k]) + "*)p[" + std::to_string(k) + "])[j];\n";
    src += "    return e;\n  }\n  __device__ __forceinline__ void store(int64_t j, const E& e) const {\n";
    for (size_t k = 0; k < nd.elt.size(); ++k)
      src += "    ((" + std::string(sc_storetype(nd.elt[k])) + "*)p[" + std::to_string(k) + "])[j] = e.f" + std::to_string(k) + ";\n";
    src += "  }\n  __device__ __forceinline__ void cas_update(int64_t j, const E& v, const Fn& fn) const {\n";
    if (nd.cas)
      src += std::string("    accjit::cas_update<E>((") + (sc_size(nd.elt[0]) == 4 ? "uint32_t" : "uint64_t") +
             "*)p[0] + j, v, fn);\n";
    src += "  }\n};\n";
    src += std::string("extern \"C\" __global__ __launch_bounds__(256) void acc_jit_permute(Args a) {\n  permute<E, Fn, Ld, Proj, Out, ") +
           (nd.cas ? "true" : "false") + ">(a, " + FN + ", " + LD + ", Proj{a.fv}, Out{" + ptr_list("a.out") + "});\n}\n";
  }
  // foreign scalar functions, at file scope ahead of the functors that call them
  std::string defs;
  for (const auto& kv : *proto.foreign) defs += kv.second;
  const std::string head = "using namespace accjit;\n";
  const size_t at = src.find(head) + head.size();
  nd.source = src.substr(0, at) + defs + src.substr(at);
  nd.key = node_key(src);
}

// ------------------------------------------------------- compile + cache
static std::string cache_dir() {
  if (const char* d = getenv("ACC_HIP_CACHE_DIR")) return d;
  std::string base;
  if (const char* x = getenv("XDG_CACHE_HOME")) base = x;
  else if (const char* h = getenv("HOME")) base = std::string(h) + "/.cache";
  else base = "/tmp";
  return base + "/accelerate-hip";
}

static void mkdirs(const std::string& d) {
  std::string p;
  for (size_t i = 0; i < d.size(); ++i) {
    p += d[i];
    if ((d[i] == '/' && i) || i + 1 == d.size()) (void)mkdir(p.c_str(), 0755);
  }
}

static bool read_file(const std::string& path, std::vector<char>& out) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  fseek(f, 0, SEEK_END);
  const long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  out.resize(n > 0 ? (size_t)n : 0);
  const bool ok = n > 0 && fread(out.data(), 1, (size_t)n, f) == (size_t)n;
  fclose(f);
  return ok;
}

static void write_file_atomic(const std::string& path, const std::vector<char>& data) {
  const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
  FILE* f = fopen(tmp.c_str(), "wb");
  if (!f) return;
  const bool ok = fwrite(data.data(), 1, data.size(), f) == data.size();
  fclose(f);
  if (ok) (void)rename(tmp.c_str(), path.c_str());
  else (void)unlink(tmp.c_str());
}

static std::mutex g_rtc_mu;  // one hiprtc compilation at a time

static int rtc_compile(const Node& nd, std::vector<char>& code, double* ms) {
  std::lock_guard<std::mutex> lk(g_rtc_mu);
  const auto t0 = std::chrono::steady_clock::now();
  const char* hdr[kNumHeaders];
  const char* names[kNumHeaders];
  for (int i = 0; i < kNumHeaders; ++i) hdr[i] = kHeaders[i].text, names[i] = kHeaders[i].name;
  hiprtcProgram p;
  if (hiprtcCreateProgram(&p, nd.source.c_str(), "acc_jit_node.hip", kNumHeaders, hdr, names) != HIPRTC_SUCCESS)
    return fail(ACC_ERROR_HIP, "hiprtcCreateProgram failed");
  const hiprtcResult r = hiprtcCompileProgram(p, kNumOptions, kOptions);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(p, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(p, &log[0]);
    hiprtcDestroyProgram(&p);
    return fail(ACC_ERROR_INVALID_VALUE, "kernel compilation failed: %.3800s", log.c_str());
  }
  size_t n = 0;
  hiprtcGetCodeSize(p, &n);
  code.resize(n);
  hiprtcGetCode(p, code.data());
  hiprtcDestroyProgram(&p);
  *ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return ACC_SUCCESS;
}

// The node's code object: from the disk cache, else compiled and stored.
static int code_object(const Node& nd, std::vector<char>& code, int* origin, double* ms) {
  const std::string dir = cache_dir();
  const std::string path = dir + "/" + nd.key + ".gfx950.co";
  *ms = 0;
  if (read_file(path, code)) {
    *origin = ACC_KERNEL_FROM_DISK_CACHE;
    return ACC_SUCCESS;
  }
  ACC_CHECK(rtc_compile(nd, code, ms));
  mkdirs(dir);
  write_file_atomic(path, code);
  *origin = ACC_KERNEL_COMPILED;
  return ACC_SUCCESS;
}

// A loaded module and its kernels on one device, shared by every handle of
// the same node (the reference's link cache, LLVM/Link/Cache.hs:50-115).
struct Loaded {
  int device = 0;
  hipModule_t mod = nullptr;
  std::map<std::string, hipFunction_t> fn;
  std::map<std::string, int> occ;  // resident workgroups per CU
  ~Loaded() {
    if (mod) {
      (void)hipSetDevice(device);
      (void)hipModuleUnload(mod);
    }
  }
};

static std::mutex g_table_mu;
static std::map<std::pair<int, std::string>, std::weak_ptr<Loaded>> g_table;

static int load_module(int device, const Node& nd, const std::vector<char>& code, std::shared_ptr<Loaded>* out) {
  ACC_CHECK(set_device(device));
  auto L = std::make_shared<Loaded>();
  L->device = device;
  ACC_HIP_CHECK(hipModuleLoadData(&L->mod, code.data()));
  std::vector<const char*> names;
  switch (nd.kind) {
    case K_FOLD: names = {"acc_jit_fold_all", "acc_jit_fold_seg_wave", "acc_jit_fold_seg_block"}; break;
    case K_FOLDSEG: names = {"acc_jit_fold_seg_wave", "acc_jit_fold_seg_block"}; break;
    case K_SCAN: names = {"acc_jit_scan", "acc_jit_scan_rows"}; break;
    case K_PERMUTE: names = {"acc_jit_permute"}; break;
    case K_STENCIL: names = {"acc_jit_stencil"}; break;
    case K_GENERATE: names = {"acc_jit_generate"}; break;
  }
  for (const char* nm : names) {
    hipFunction_t f;
    ACC_HIP_CHECK(hipModuleGetFunction(&f, L->mod, nm));
    L->fn[nm] = f;
    int nb = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 256, 0) != hipSuccess || nb < 1) {
      (void)hipGetLastError();
      nb = 1;
    }
    L->occ[nm] = nb;
  }
  *out = L;
  return ACC_SUCCESS;
}

}  // namespace jit
}  // namespace acc

using namespace acc;
using namespace acc::jit;

struct acc_kernel_s {
  int device = 0;
  Node node;
  std::shared_ptr<Loaded> mod;
  std::vector<char> image;
  int origin = 0;
  double compile_ms = 0;
};

namespace acc {
namespace jit {

static const char kImageMagic[8] = {'A', 'C', 'C', 'J', 'I', 'T', '1', '\0'};

static int make_kernel(int device, const char* text, const void* image, size_t bytes, acc_kernel_t* out) {
  if (!text || !out) return fail(ACC_ERROR_INVALID_VALUE, "kernel: null argument");
  auto k = std::make_unique<acc_kernel_s>();
  k->device = device;
  try {
    build_node(text, k->node);
  } catch (const Error& e) {
    return fail(ACC_ERROR_INVALID_VALUE, "kernel node: %s", e.what());
  }
  const std::pair<int, std::string> id(device, k->node.key);
  {
    std::lock_guard<std::mutex> lk(g_table_mu);
    auto it = g_table.find(id);
    if (it != g_table.end()) k->mod = it->second.lock();
  }
  if (image) {
    // an embedded image: header (magic, key) + code object
    if (bytes < 8 + 32 || memcmp(image, kImageMagic, 8) != 0)
      return fail(ACC_ERROR_INVALID_VALUE, "kernel image: not an image made by acc_hip_kernel_image");
    if (memcmp((const char*)image + 8, k->node.key.data(), 32) != 0)
      return fail(ACC_ERROR_INVALID_VALUE, "kernel image: made for another node or library version");
    k->image.assign((const char*)image + 40, (const char*)image + bytes);
    k->origin = ACC_KERNEL_FROM_IMAGE;
  } else if (k->mod) {
    k->origin = ACC_KERNEL_LOADED;
  } else {
    ACC_CHECK(code_object(k->node, k->image, &k->origin, &k->compile_ms));
  }
  if (!k->mod) {
    std::shared_ptr<Loaded> L;
    ACC_CHECK(load_module(device, k->node, k->image, &L));
    std::lock_guard<std::mutex> lk(g_table_mu);
    auto& slot = g_table[id];
    if (auto cur = slot.lock()) k->mod = cur;  // another thread loaded it meanwhile
    else slot = k->mod = L;
  }
  *out = k.release();
  return ACC_SUCCESS;
}

static int launch(acc_kernel_t k, const char* name, unsigned grid, acc_stream_t s, accjit::Args& a) {
  if (!grid) return ACC_SUCCESS;  // zero-size launches are skipped (PTX/Execute.hs:847-854)
  hipFunction_t f = k->mod->fn.at(name);
  void* params[] = {&a};
  ACC_HIP_CHECK(hipModuleLaunchKernel(f, grid, 1, 1, 256, 1, 1, 0, s->stream, params, nullptr));
  return ACC_SUCCESS;
}

static bool aligned_fields(const void* const* p, const std::vector<int>& f) {
  for (size_t k = 0; k < f.size(); ++k)
    if ((uintptr_t)p[k] % (uintptr_t)(4 * sc_size(f[k]))) return false;
  return true;
}

static int check_fields(const void* const* p, const std::vector<int>& f, const char* what) {
  for (size_t k = 0; k < f.size(); ++k) {
    if (!p[k]) return fail(ACC_ERROR_INVALID_VALUE, "%s: field %zu is null", what, k);
    if ((uintptr_t)p[k] % (uintptr_t)sc_size(f[k]))
      return fail(ACC_ERROR_INVALID_VALUE, "%s: field %zu is not aligned to its element", what, k);
  }
  return ACC_SUCCESS;
}

static int run(acc_kernel_t k, acc_stream_t s, const acc_kernel_args* ka) {
  const Node& nd = k->node;
  if (s->device != k->device) return fail(ACC_ERROR_INVALID_VALUE, "kernel: stream on another device");
  if (ka->outer < 0 || ka->inner < 0 || ka->nseg < 0) return fail(ACC_ERROR_INVALID_VALUE, "kernel: negative extent");
  if (ka->nfree < 0 || ka->nfree > accjit::kMaxFree || (size_t)ka->nfree < nd.free.size())
    return fail(ACC_ERROR_INVALID_VALUE, "kernel: %zu free arrays expected, %d given", nd.free.size(), ka->nfree);
  ACC_CHECK(set_device(s->device));
  const DeviceInfo* di;
  ACC_CHECK(device_info(s->device, &di));
  const int64_t cus = di->compute_units;
  accjit::Args a;
  memset(&a, 0, sizeof a);
  for (int f = 0; f < accjit::kMaxFields; ++f) {
    a.in[f] = ka->in[f];
    a.out[f] = ka->out[f];
    a.sums[f] = ka->sums[f];
  }
  a.offsets = ka->offsets;
  a.outer = ka->outer;
  a.inner = ka->inner;
  a.nseg = ka->nseg;
  for (size_t f = 0; f < nd.free.size(); ++f) {
    if (!ka->free_ptr[f] && ka->free_len[f] > 0) return fail(ACC_ERROR_INVALID_VALUE, "kernel: free array %zu is null", f);
    a.fv.p[f] = ka->free_ptr[f];
    a.fv.n[f] = ka->free_len[f];
  }
  const size_t NWm = (size_t)(nd.lay.maybe_size + 3) / 4, NWe = (size_t)(nd.lay.size + 3) / 4;
  const int64_t n = ka->outer * ka->inner;
  const bool in_ok = nd.delayed || nd.kind == K_GENERATE || nd.kind == K_STENCIL;

  switch (nd.kind) {
    case K_FOLD: {
      // fold / fold1 (PTX/Execute.hs:222-340): out[r] = seed (+) fold of row r
      if (ka->outer == 0) return ACC_SUCCESS;
      if (ka->inner == 0 && !nd.has_seed) return fail(ACC_ERROR_EMPTY_ARRAY, "fold1: empty array");
      if (!in_ok && ka->inner) ACC_CHECK(check_fields(ka->in, nd.elt, "fold input"));
      ACC_CHECK(check_fields((const void* const*)ka->out, nd.elt, "fold output"));
      a.vec = (in_ok || aligned_fields(ka->in, nd.elt)) ? 1u : 0u;
      if (ka->outer == 1 && ka->inner > 0) {
        int64_t nb = (ka->inner + 4095) / 4096;
        if (nb > cus * 8) nb = cus * 8;
        int64_t chunk = (ka->inner + nb - 1) / nb;
        chunk = (chunk + 1023) / 1024 * 1024;
        nb = (ka->inner + chunk - 1) / chunk;
        a.chunk = chunk;
        acc::SyncCtl* ctl;
        uint64_t* st;
        ACC_CHECK(stream_workspace(s, (size_t)nb * NWm * 8, &ctl, &st));
        a.ctl = ctl;
        a.status = st;
        return launch(k, "acc_jit_fold_all", (unsigned)nb, s, a);
      }
      a.nseg = ka->outer;  // one uniform segment per row: rows reduced as segments of a single row
      a.outer = 1;
      a.inner = n;
      const bool wave = ka->inner < 512;
      const int64_t work = wave ? (ka->outer + 3) / 4 : ka->outer;
      const int64_t grid = work < cus * 16 ? work : cus * 16;
      // segments of `inner` elements: OffUniform uses a.inner as the segment length
      a.inner = ka->inner;
      return launch(k, wave ? "acc_jit_fold_seg_wave" : "acc_jit_fold_seg_block", (unsigned)grid, s, a);
    }
    case K_FOLDSEG: {
      // foldSeg / fold1Seg (PTX/Execute.hs:343-374)
      if (ka->outer == 0 || ka->nseg == 0) return ACC_SUCCESS;
      if (!ka->offsets) return fail(ACC_ERROR_INVALID_VALUE, "foldSeg: null offsets");
      if (!in_ok && ka->inner) ACC_CHECK(check_fields(ka->in, nd.elt, "foldSeg input"));
      ACC_CHECK(check_fields((const void* const*)ka->out, nd.elt, "foldSeg output"));
      // a warp per segment when the mean segment is short (PTX/Execute.hs:361-366)
      const bool wave = ka->inner / ka->nseg < 2 * 256;
      const int64_t segs = ka->outer * ka->nseg;
      const int64_t work = wave ? (segs + 3) / 4 : segs;
      const int64_t grid = work < cus * 16 ? work : cus * 16;
      return launch(k, wave ? "acc_jit_fold_seg_wave" : "acc_jit_fold_seg_block", (unsigned)grid, s, a);
    }
    case K_SCAN: {
      // scanl/scanl1/scanr/scanr1 (PTX/Execute.hs:377-488) and scan' (:491-594)
      if (ka->outer == 0) return ACC_SUCCESS;
      const bool seed_last = nd.has_seed && !nd.prime && nd.rev;  // scanr z: out[n] = z
      const bool excl = nd.prime || (nd.has_seed && !nd.rev);      // scanl z, scan'
      a.exclusive = excl ? 1u : 0u;
      a.has_total = nd.prime || (nd.has_seed && !nd.rev) ? 1u : (seed_last ? 2u : 0u);
      a.out_row = nd.has_seed && !nd.prime ? ka->inner + 1 : ka->inner;
      a.chunk = ka->inner;  // Tot: the total / seed goes to index `inner` of the row
      if (ka->inner == 0) {
        if (!nd.has_seed) return ACC_SUCCESS;
        // empty rows: scanl z = [z], scan' sums = z (PTX/Execute.hs:388, 500-509)
        a.exclusive = 0;
        a.has_total = 2;
      }
      ACC_CHECK(check_fields((const void* const*)ka->out, nd.elt, "scan output"));
      if (nd.prime) ACC_CHECK(check_fields((const void* const*)ka->sums, nd.elt, "scan' sums"));
      if (!in_ok && ka->inner) ACC_CHECK(check_fields(ka->in, nd.elt, "scan input"));
      a.vec = ((in_ok || aligned_fields(ka->in, nd.elt)) && aligned_fields((const void* const*)ka->out, nd.elt)) ? 1u : 0u;
      if (ka->outer == 1 && ka->inner > 0) {
        const int64_t TILE = 256LL * nd.K * 4;
        const int64_t nt = (ka->inner + TILE - 1) / TILE;
        if (nt > 0x7fffffffLL) return fail(ACC_ERROR_UNSUPPORTED, "scan: too many elements");
        a.ntiles = (uint32_t)nt;
        acc::SyncCtl* ctl;
        uint64_t* st;
        ACC_CHECK(stream_workspace(s, (size_t)nt * NWe * 8, &ctl, &st));
        a.ctl = ctl;
        a.status = st;
        int64_t grid = cus * k->mod->occ.at("acc_jit_scan");
        if (grid > nt) grid = nt;
        return launch(k, "acc_jit_scan", (unsigned)grid, s, a);
      }
      const int64_t grid = ka->outer < cus * 8 ? ka->outer : cus * 8;
      return launch(k, "acc_jit_scan_rows", (unsigned)grid, s, a);
    }
    case K_PERMUTE: {
      // permute (PTX/Execute.hs:597-649): clone the defaults unless in place
      const int64_t m = ka->nseg;
      ACC_CHECK(check_fields((const void* const*)ka->out, nd.elt, "permute output"));
      if (ka->defaults[0] && ka->defaults[0] != ka->out[0]) {
        ACC_CHECK(check_fields(ka->defaults, nd.elt, "permute defaults"));
        for (size_t f = 0; f < nd.elt.size(); ++f)
          ACC_CHECK(acc_hip_memcpy_d2d_async(s, ka->out[f], ka->defaults[f], (size_t)m * sc_size(nd.elt[f])));
      }
      if (n == 0 || m == 0) return ACC_SUCCESS;
      if (!in_ok) ACC_CHECK(check_fields(ka->in, nd.elt, "permute source"));
      if (!nd.cas) {
        // the locks are untagged words: stream scratch, not the status region
        void* lk;
        ACC_CHECK(stream_scratch(s, (size_t)m * 4, &lk));
        a.locks = (uint32_t*)lk;
        const uint32_t zero = 0;
        // a fill kernel, not a memset node (kept capturable)
        ACC_CHECK(acc_hip_fill_async(s, a.locks, &zero, 4, (size_t)m));
      }
      int64_t grid = (n + 255) / 256;
      if (grid > cus * 16) grid = cus * 16;
      return launch(k, "acc_jit_permute", (unsigned)grid, s, a);
    }
    case K_STENCIL: {
      // stencil1 on DIM2 (PTX/Execute.hs:652-776): outer = height, inner = width
      if (n == 0) return ACC_SUCCESS;
      if (!ka->in[0] || !ka->out[0]) return fail(ACC_ERROR_INVALID_VALUE, "stencil: null array");
      if (ka->in[0] == ka->out[0]) return fail(ACC_ERROR_INVALID_VALUE, "stencil: output must not alias the input");
      const int64_t tiles = ((ka->outer + 15) / 16) * ((ka->inner + 63) / 64);
      if (tiles > 0x7fffffffLL) return fail(ACC_ERROR_UNSUPPORTED, "stencil: too large");
      return launch(k, "acc_jit_stencil", (unsigned)tiles, s, a);
    }
    case K_GENERATE: {
      if (n == 0) return ACC_SUCCESS;
      ACC_CHECK(check_fields((const void* const*)ka->out, nd.elt, "generate output"));
      a.vec = aligned_fields((const void* const*)ka->out, nd.elt) ? 1u : 0u;
      int64_t grid = (n + 1023) / 1024;
      if (grid > cus * 16) grid = cus * 16;
      return launch(k, "acc_jit_generate", (unsigned)grid, s, a);
    }
  }
  return fail(ACC_ERROR_UNSUPPORTED, "kernel: unknown skeleton");
}

}  // namespace jit
}  // namespace acc

extern "C" {

int acc_hip_node_source(const char* node, char* buf, size_t cap, size_t* len) {
  if (!node) return fail(ACC_ERROR_INVALID_VALUE, "null node");
  Node nd;
  try {
    build_node(node, nd);
  } catch (const Error& e) {
    return fail(ACC_ERROR_INVALID_VALUE, "kernel node: %s", e.what());
  }
  if (len) *len = nd.source.size();
  if (buf && cap) {
    const size_t n = nd.source.size() < cap - 1 ? nd.source.size() : cap - 1;
    memcpy(buf, nd.source.data(), n);
    buf[n] = '\0';
  }
  return ACC_SUCCESS;
}

int acc_hip_node_precompile(const char* node, int* origin) {
  if (!node) return fail(ACC_ERROR_INVALID_VALUE, "null node");
  Node nd;
  try {
    build_node(node, nd);
  } catch (const Error& e) {
    return fail(ACC_ERROR_INVALID_VALUE, "kernel node: %s", e.what());
  }
  std::vector<char> code;
  int o = 0;
  double ms = 0;
  ACC_CHECK(code_object(nd, code, &o, &ms));
  if (origin) *origin = o;
  return ACC_SUCCESS;
}

int acc_hip_kernel_compile(int device, const char* node, acc_kernel_t* k) {
  return make_kernel(device, node, nullptr, 0, k);
}

int acc_hip_kernel_from_image(int device, const char* node, const void* image, size_t bytes, acc_kernel_t* k) {
  if (!image) return fail(ACC_ERROR_INVALID_VALUE, "kernel image: null");
  return make_kernel(device, node, image, bytes, k);
}

int acc_hip_kernel_image(acc_kernel_t k, void* buf, size_t cap, size_t* bytes) {
  if (!k || !bytes) return fail(ACC_ERROR_INVALID_VALUE, "kernel image: null argument");
  if (k->image.empty()) {
    // the module came from the in-process table: fetch its code object again
    int o;
    double ms;
    ACC_CHECK(code_object(k->node, k->image, &o, &ms));
  }
  const size_t need = 40 + k->image.size();
  *bytes = need;
  if (!buf) return ACC_SUCCESS;
  if (cap < need) return fail(ACC_ERROR_INVALID_VALUE, "kernel image: %zu bytes needed", need);
  memcpy(buf, kImageMagic, 8);
  memcpy((char*)buf + 8, k->node.key.data(), 32);
  memcpy((char*)buf + 40, k->image.data(), k->image.size());
  return ACC_SUCCESS;
}

int acc_hip_kernel_info(acc_kernel_t k, acc_kernel_info* info) {
  if (!k || !info) return fail(ACC_ERROR_INVALID_VALUE, "kernel info: null argument");
  memset(info, 0, sizeof *info);
  info->origin = k->origin;
  info->compile_ms = k->compile_ms;
  snprintf(info->key, sizeof info->key, "%s", k->node.key.c_str());
  info->lock_free_permute = k->node.kind == K_PERMUTE ? (k->node.cas ? 1 : 0) : -1;
  return ACC_SUCCESS;
}

int acc_hip_kernel_run(acc_kernel_t k, acc_stream_t s, const acc_kernel_args* args) {
  if (!k || !s || !args) return fail(ACC_ERROR_INVALID_VALUE, "kernel run: null argument");
  return run(k, s, args);
}

int acc_hip_kernel_destroy(acc_kernel_t k) {
  delete k;
  return ACC_SUCCESS;
}

}  // extern "C"
