// Expression language of the run-time compiled skeletons: parser, type
// checker and emitter (declarations and the language summary in
// jit_expr.hpp; the semantics of each primitive in jit/prelude.hpp).
#include "jit_expr.hpp"

#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <map>

namespace acc {
namespace jit {

static const char* kName[NSC] = {"i8", "i16", "i32", "i64", "u8", "u16", "u32", "u64", "f32", "f64", "bool"};
static const char* kCType[NSC] = {"int8_t",   "int16_t",  "int32_t", "int64_t", "uint8_t", "uint16_t",
                                  "uint32_t", "uint64_t", "float",   "double",  "bool"};
static const char* kStore[NSC] = {"int8_t",   "int16_t",  "int32_t", "int64_t", "uint8_t", "uint16_t",
                                  "uint32_t", "uint64_t", "float",   "double",  "uint8_t"};
static const int kSize[NSC] = {1, 2, 4, 8, 1, 2, 4, 8, 4, 8, 1};

const char* sc_name(int t) { return t >= 0 && t < NSC ? kName[t] : "?"; }
const char* sc_ctype(int t) { return kCType[t]; }
const char* sc_storetype(int t) { return kStore[t]; }
int sc_size(int t) { return kSize[t]; }
int sc_parse(const std::string& s) {
  for (int i = 0; i < NSC; ++i)
    if (s == kName[i]) return i;
  // Accelerate's names for the same types
  if (s == "int") return I64;
  if (s == "word") return U64;
  if (s == "float") return F32;
  if (s == "double") return F64;
  return -1;
}

[[noreturn]] static void err(const std::string& m) { throw Error(m); }

std::string SX::str() const {
  if (atom && quoted) {
    std::string q = "\"";
    for (char c : a) q += (c == '"' || c == '\\') ? std::string("\\") + c : std::string(1, c);
    return q + "\"";
  }
  if (atom) return a;
  std::string s = "(";
  for (size_t i = 0; i < l.size(); ++i) s += (i ? " " : "") + l[i].str();
  return s + ")";
}

// ------------------------------------------------------------------ parser
SX parse_sx(const std::string& text) {
  size_t i = 0;
  std::vector<SX> stack(1);
  auto skip = [&]() {
    while (i < text.size()) {
      if (isspace((unsigned char)text[i])) ++i;
      else if (text[i] == ';') { while (i < text.size() && text[i] != '\n') ++i; }
      else break;
    }
  };
  while (true) {
    skip();
    if (i >= text.size()) break;
    const char c = text[i];
    if (c == '(') {
      stack.emplace_back();
      ++i;
    } else if (c == '"') {
      // a string literal: \" and \\ escape, anything else is kept as is
      SX a;
      a.atom = true;
      a.quoted = true;
      size_t j = i + 1;
      for (; j < text.size() && text[j] != '"'; ++j) {
        if (text[j] == '\\' && j + 1 < text.size() && (text[j + 1] == '"' || text[j + 1] == '\\')) ++j;
        a.a += text[j];
      }
      if (j >= text.size()) err("unterminated string literal");
      stack.back().l.push_back(std::move(a));
      i = j + 1;
    } else if (c == ')') {
      if (stack.size() < 2) err("unbalanced ')'");
      SX done = std::move(stack.back());
      stack.pop_back();
      stack.back().l.push_back(std::move(done));
      ++i;
    } else {
      size_t j = i;
      while (j < text.size() && !isspace((unsigned char)text[j]) && text[j] != '(' && text[j] != ')' && text[j] != ';') ++j;
      SX a;
      a.atom = true;
      a.a = text.substr(i, j - i);
      stack.back().l.push_back(std::move(a));
      i = j;
    }
  }
  if (stack.size() != 1) err("unbalanced '('");
  if (stack[0].l.size() != 1) err("expected exactly one expression");
  return std::move(stack[0].l[0]);
}

// ----------------------------------------------------------------- layout
Layout layout_of(const std::vector<int>& f) {
  Layout L;
  int off = 0;
  for (int t : f) {
    const int s = kSize[t];
    off = (off + s - 1) / s * s + s;
    if (s > L.align) L.align = s;
  }
  L.size = (off + L.align - 1) / L.align * L.align;
  const int ma = L.align > 4 ? L.align : 4;
  const int ok = (L.size + 3) / 4 * 4 + 4;
  L.maybe_size = (ok + ma - 1) / ma * ma;
  return L;
}

std::string struct_decl(const char* name, const std::vector<int>& f) {
  std::string s = std::string("struct ") + name + " {";
  for (size_t k = 0; k < f.size(); ++k) s += std::string(" ") + kStore[f[k]] + " f" + std::to_string(k) + ";";
  return s + " };\n";
}

// ---------------------------------------------------------------- emitter
std::string Emitter::bind(int t, const std::string& expr) {
  const std::string n = "t" + std::to_string(tmp++);
  body += "    const " + std::string(kCType[t]) + " " + n + " = " + expr + ";\n";
  return n;
}

static Val scalar(int t, const std::string& v) { return Val{{v}, {t}}; }

static int want_scalar(const Val& v, const char* what) {
  if (v.t.size() != 1) err(std::string(what) + ": expected a scalar, got a tuple of " + std::to_string(v.t.size()));
  return v.t[0];
}

static std::string hex64(uint64_t x) {
  char b[32];
  snprintf(b, sizeof b, "0x%llxull", (unsigned long long)x);
  return b;
}

static std::string literal(int t, const std::string& s) {
  if (t == BOOL) {
    if (s == "true" || s == "1") return "true";
    if (s == "false" || s == "0") return "false";
    err("bad Bool literal " + s);
  }
  char* end = nullptr;
  if (sc_float(t)) {
    errno = 0;
    const double d = strtod(s.c_str(), &end);
    if (end == s.c_str() || *end) err("bad floating literal " + s);
    if (t == F32) {
      const float f = (float)d;
      uint32_t b;
      memcpy(&b, &f, 4);
      return "ap::bitcast<float>((uint32_t)" + hex64(b) + ")";
    }
    uint64_t b;
    memcpy(&b, &d, 8);
    return "ap::bitcast<double>((uint64_t)" + hex64(b) + ")";
  }
  uint64_t bits;
  if (sc_signed(t)) {
    errno = 0;
    const long long v = strtoll(s.c_str(), &end, 0);
    if (end == s.c_str() || *end || errno) err("bad integer literal " + s);
    bits = (uint64_t)v;
  } else {
    errno = 0;
    const unsigned long long v = strtoull(s.c_str(), &end, 0);
    if (end == s.c_str() || *end || errno || s[0] == '-') err("bad unsigned literal " + s);
    bits = v;
  }
  // the bit pattern, narrowed to the type (an out-of-range literal wraps)
  return std::string("((") + kCType[t] + ")(" + hex64(bits) + "))";
}

static int type_arg(const SX& e, size_t k, const char* op) {
  if (e.l.size() <= k || !e.l[k].atom) err(std::string(op) + ": expected a type name");
  const int t = sc_parse(e.l[k].a);
  if (t < 0) err(std::string(op) + ": unknown type " + e.l[k].a);
  return t;
}

Val Emitter::emit(const SX& e) {
  if (e.atom) {
    for (auto it = env.rbegin(); it != env.rend(); ++it)
      if (it->first == e.a) return it->second;
    err("unbound variable " + e.a);
  }
  if (e.l.empty() || !e.l[0].atom) err("expected (operator args...), got " + e.str());
  const std::string& op = e.l[0].a;
  const size_t na = e.l.size() - 1;
  auto arity = [&](size_t k) {
    if (na != k) err(op + ": expected " + std::to_string(k) + " arguments, got " + std::to_string(na));
  };

  if (op == "const") {
    arity(2);
    const int t = type_arg(e, 1, "const");
    if (!e.l[2].atom) err("const: expected a literal");
    return scalar(t, literal(t, e.l[2].a));
  }
  if (op == "let") {
    if (na != 2 || e.l[1].atom) err("let: expected (let ((x e) ...) body)");
    const size_t mark = env.size();
    for (const SX& b : e.l[1].l) {
      if (b.atom || b.l.size() != 2 || !b.l[0].atom) err("let: bad binding " + b.str());
      Val v = emit(b.l[1]);
      Val named;
      named.t = v.t;
      for (size_t k = 0; k < v.v.size(); ++k) named.v.push_back(bind(v.t[k], v.v[k]));
      env.emplace_back(b.l[0].a, named);
    }
    Val r = emit(e.l[2]);
    env.resize(mark);
    return r;
  }
  if (op == "cond") {
    arity(3);
    const Val c = emit(e.l[1]);
    if (want_scalar(c, "cond") != BOOL) err("cond: the condition must be Bool");
    const Val a = emit(e.l[2]), b = emit(e.l[3]);
    if (a.t != b.t) err("cond: branches of different types");
    Val r;
    r.t = a.t;
    for (size_t k = 0; k < a.v.size(); ++k) r.v.push_back(bind(a.t[k], c.v[0] + " ? " + a.v[k] + " : " + b.v[k]));
    return r;
  }
  if (op == "tuple") {
    Val r;
    for (size_t k = 1; k < e.l.size(); ++k) {
      const Val v = emit(e.l[k]);
      r.v.insert(r.v.end(), v.v.begin(), v.v.end());
      r.t.insert(r.t.end(), v.t.begin(), v.t.end());
    }
    return r;
  }
  if (op == "prj") {
    arity(2);
    if (!e.l[1].atom) err("prj: expected a field index");
    const int k = atoi(e.l[1].a.c_str());
    const Val v = emit(e.l[2]);
    if (k < 0 || k >= (int)v.v.size()) err("prj: field " + e.l[1].a + " of a tuple of " + std::to_string(v.v.size()));
    return scalar(v.t[k], v.v[k]);
  }
  if (op == "read") {
    arity(2);
    if (!e.l[1].atom) err("read: expected a free array index");
    const int k = atoi(e.l[1].a.c_str());
    if (k < 0 || k >= (int)free_types.size()) err("read: no free array " + e.l[1].a);
    const Val i = emit(e.l[2]);
    if (want_scalar(i, "read") != I64) err("read: the index must be Int (i64)");
    const int t = free_types[k];
    const std::string rd = std::string("ap::read<") + kStore[t] + ">(fv.p[" + std::to_string(k) + "], fv.n[" +
                           std::to_string(k) + "], " + i.v[0] + ")";
    return scalar(t, bind(t, t == BOOL ? "(" + rd + " != 0)" : rd));
  }
  if (op == "foreign") {
    // (foreign "BODY" (fun (x ..) FALLBACK) E ..): a foreign scalar function
    // (ForeignExp, PTX/Foreign.hs:74-84): BODY is the body of a device
    // function of the parameters x .. (typed by the arguments E ..), compiled
    // into the kernel; the fallback gives the function's meaning where no
    // foreign code runs and, type-checked on the same arguments, its result
    // type.  Numeric scalars only.
    if (na < 2 || !e.l[1].atom || !e.l[1].quoted) err("foreign: expected (foreign \"body\" (fun (x ..) e) args ..)");
    const SX& fb = e.l[2];
    if (!fb.is("fun") || fb.l.size() != 3 || fb.l[1].atom) err("foreign: the fallback must be (fun (x ..) e)");
    const std::vector<SX>& ps = fb.l[1].l;
    if (ps.size() != na - 2) err("foreign: the fallback takes " + std::to_string(ps.size()) + " arguments, given " +
                                 std::to_string(na - 2));
    if (!foreign) err("foreign: not available here");
    std::vector<Val> args;
    std::vector<int> ts;
    for (size_t k = 3; k < e.l.size(); ++k) {
      args.push_back(emit(e.l[k]));
      ts.push_back(want_scalar(args.back(), "foreign"));
      if (!sc_num(ts.back())) err("foreign: numeric scalar arguments only");
    }
    // the fallback, typed on placeholders of the arguments' types (closed:
    // only its parameters and the free arrays are in scope)
    Emitter chk;
    chk.free_types = free_types;
    chk.foreign = foreign;
    for (size_t k = 0; k < ps.size(); ++k) {
      if (!ps[k].atom) err("foreign: parameter names must be symbols");
      Val v;
      v.t = {ts[k]};
      v.v = {std::string("((") + kCType[ts[k]] + ")0)"};
      chk.env.emplace_back(ps[k].a, v);
    }
    const Val fr = chk.emit(fb.l[2]);
    const int rt = want_scalar(fr, "foreign fallback");
    if (!sc_num(rt)) err("foreign: numeric scalar result only");
    std::string sig = std::string(kCType[rt]) + "(";
    for (size_t k = 0; k < ts.size(); ++k) sig += std::string(k ? ", " : "") + kCType[ts[k]] + " " + ps[k].a;
    sig += ")";
    unsigned long long h = 0xcbf29ce484222325ull;
    for (unsigned char c : sig + "\n" + e.l[1].a) { h ^= c; h *= 0x100000001b3ull; }
    char name[40];
    snprintf(name, sizeof name, "acc_foreign_%016llx", h);
    (*foreign)[name] = std::string("static __device__ __forceinline__ ") + kCType[rt] + " " + name +
                       sig.substr(sig.find('(')) + " {\n" + e.l[1].a + "\n}\n";
    std::string call = std::string(name) + "(";
    for (size_t k = 0; k < args.size(); ++k) call += (k ? ", " : "") + args[k].v[0];
    call += ")";
    return scalar(rt, bind(rt, call));
  }
  if (op == "nb") {
    arity(2);
    if (!stencil) err("nb: only in a stencil function");
    if (!e.l[1].atom || !e.l[2].atom) err("nb: expected literal offsets (nb dy dx)");
    const int dy = atoi(e.l[1].a.c_str()), dx = atoi(e.l[2].a.c_str());
    if (dy < -nb_ry || dy > nb_ry || dx < -nb_rx || dx > nb_rx) err("nb: offset outside the stencil");
    const std::string rd = "nb[" + std::to_string(dy * nb_stride + dx) + "]";
    return scalar(nb_type, nb_type == BOOL ? "(" + rd + " != 0)" : rd);
  }

  // ---- primitives
  std::vector<Val> args;
  const bool conv = op == "fromintegral" || op == "tofloating" || op == "bitcast" || op == "truncate" ||
                    op == "round" || op == "floor" || op == "ceiling";
  int target = -1;
  if (conv) {
    arity(2);
    target = type_arg(e, 1, op.c_str());
    args.push_back(emit(e.l[2]));
  } else {
    for (size_t k = 1; k < e.l.size(); ++k) args.push_back(emit(e.l[k]));
  }
  std::vector<int> ts;
  for (auto& a : args) ts.push_back(want_scalar(a, op.c_str()));
  auto A = [&](size_t k) { return args[k].v[0]; };
  auto need = [&](bool ok, const char* what) {
    if (!ok) err(op + ": " + what);
  };
  auto same2 = [&]() {
    arity(2);
    need(ts[0] == ts[1], "operands of different types");
  };
  auto call1 = [&](const char* f, int rt) { return scalar(rt, bind(rt, std::string("ap::") + f + "<" + kCType[ts[0]] + ">(" + A(0) + ")")); };
  auto call2 = [&](const char* f, int rt) {
    return scalar(rt, bind(rt, std::string("ap::") + f + "<" + kCType[ts[0]] + ">(" + A(0) + ", " + A(1) + ")"));
  };

  static const std::map<std::string, const char*> num2 = {{"add", "add"}, {"sub", "sub"}, {"mul", "mul"}};
  static const std::map<std::string, const char*> num1 = {{"neg", "neg"}, {"abs", "abs_"}, {"signum", "signum"}};
  static const std::map<std::string, const char*> int2 = {{"quot", "quot"}, {"rem", "rem"}, {"idiv", "idiv"},
                                                          {"div", "idiv"},  {"mod", "mod"}, {"band", "band"},
                                                          {"bor", "bor"},   {"bxor", "bxor"}};
  static const std::map<std::string, const char*> shifts = {
      {"shiftl", "shiftl"}, {"shiftr", "shiftr"}, {"rotatel", "rotatel"}, {"rotater", "rotater"}};
  static const std::map<std::string, const char*> bitcount = {{"popcount", "popcount"}, {"clz", "clz"}, {"ctz", "ctz"}};
  static const std::map<std::string, const char*> flt1 = {
      {"recip", "recip"}, {"sin", "sin_"},     {"cos", "cos_"},     {"tan", "tan_"},   {"sinh", "sinh_"},
      {"cosh", "cosh_"},  {"tanh", "tanh_"},   {"asin", "asin_"},   {"acos", "acos_"}, {"atan", "atan_"},
      {"asinh", "asinh_"}, {"acosh", "acosh_"}, {"atanh", "atanh_"}, {"exp", "exp_"},   {"sqrt", "sqrt_"},
      {"log", "log_"}};
  static const std::map<std::string, const char*> flt2 = {{"fdiv", "fdiv"}, {"atan2", "atan2_"}, {"fpow", "fpow"}};
  static const std::map<std::string, const char*> rel = {{"lt", "lt"},   {"gt", "gt"}, {"lte", "lte"},
                                                         {"gte", "gte"}, {"eq", "eq"}, {"neq", "neq"}};

  if (auto it = num2.find(op); it != num2.end()) { same2(); need(sc_num(ts[0]), "numeric operands"); return call2(it->second, ts[0]); }
  if (auto it = num1.find(op); it != num1.end()) { arity(1); need(sc_num(ts[0]), "a numeric operand"); return call1(it->second, ts[0]); }
  if (auto it = int2.find(op); it != int2.end()) { same2(); need(sc_int(ts[0]), "integral operands"); return call2(it->second, ts[0]); }
  if (op == "quotrem" || op == "divmod") {
    same2();
    need(sc_int(ts[0]), "integral operands");
    const bool q = op == "quotrem";
    Val r;
    r.t = {ts[0], ts[0]};
    const std::string T = kCType[ts[0]];
    r.v.push_back(bind(ts[0], "ap::" + std::string(q ? "quot" : "idiv") + "<" + T + ">(" + A(0) + ", " + A(1) + ")"));
    r.v.push_back(bind(ts[0], "ap::" + std::string(q ? "rem" : "mod") + "<" + T + ">(" + A(0) + ", " + A(1) + ")"));
    return r;
  }
  if (op == "bnot") { arity(1); need(sc_int(ts[0]), "an integral operand"); return call1("bnot", ts[0]); }
  if (auto it = shifts.find(op); it != shifts.end()) {
    arity(2);
    need(sc_int(ts[0]) && ts[1] == I64, "an integral value and an Int (i64) amount");
    return call2(it->second, ts[0]);
  }
  if (auto it = bitcount.find(op); it != bitcount.end()) {
    arity(1);
    need(sc_int(ts[0]), "an integral operand");
    return call1(it->second, I64);
  }
  if (auto it = flt1.find(op); it != flt1.end()) { arity(1); need(sc_float(ts[0]), "a floating operand"); return call1(it->second, ts[0]); }
  if (auto it = flt2.find(op); it != flt2.end()) { same2(); need(sc_float(ts[0]), "floating operands"); return call2(it->second, ts[0]); }
  if (op == "logbase") {
    same2();
    need(sc_float(ts[0]), "floating operands");
    // a constant base of 2 or 10 uses log2 / log10 (Arithmetic.hs:375-392)
    const SX& b = e.l[1];
    if (b.is("const") && b.l.size() == 3 && b.l[2].atom) {
      const double d = strtod(b.l[2].a.c_str(), nullptr);
      if (d == 2.0) return scalar(ts[0], bind(ts[0], std::string("ap::log2_<") + kCType[ts[0]] + ">(" + A(1) + ")"));
      if (d == 10.0) return scalar(ts[0], bind(ts[0], std::string("ap::log10_<") + kCType[ts[0]] + ">(" + A(1) + ")"));
    }
    return call2("logbase", ts[0]);
  }
  if (op == "isnan" || op == "isinfinite") {
    arity(1);
    need(sc_float(ts[0]), "a floating operand");
    return call1(op == "isnan" ? "isnan_" : "isinfinite", BOOL);
  }
  if (op == "truncate" || op == "round" || op == "floor" || op == "ceiling") {
    need(sc_float(ts[0]) && sc_int(target), "(op IntegralType floating-expression)");
    return scalar(target, bind(target, "ap::" + op + "_<" + kCType[target] + ", " + kCType[ts[0]] + ">(" + A(0) + ")"));
  }
  if (auto it = rel.find(op); it != rel.end()) {
    same2();
    need(sc_num(ts[0]) || ((op == "eq" || op == "neq") && ts[0] == BOOL), "comparable operands");
    return call2(it->second, BOOL);
  }
  if (op == "min" || op == "max") {
    same2();
    need(sc_num(ts[0]), "numeric operands");
    return call2(op == "min" ? "min_" : "max_", ts[0]);
  }
  if (op == "land" || op == "lor") {
    same2();
    need(ts[0] == BOOL, "Bool operands");
    return scalar(BOOL, bind(BOOL, "(" + A(0) + (op == "land" ? " && " : " || ") + A(1) + ")"));
  }
  if (op == "lnot") {
    arity(1);
    need(ts[0] == BOOL, "a Bool operand");
    return scalar(BOOL, bind(BOOL, "(!" + A(0) + ")"));
  }
  if (op == "fromintegral") {
    need(sc_int(ts[0]) && sc_num(target), "(fromintegral NumType integral-expression)");
    return scalar(target, bind(target, "ap::fromintegral<" + std::string(kCType[target]) + ", " + kCType[ts[0]] + ">(" + A(0) + ")"));
  }
  if (op == "tofloating") {
    need(sc_num(ts[0]) && sc_float(target), "(tofloating FloatingType numeric-expression)");
    return scalar(target, bind(target, "ap::tofloating<" + std::string(kCType[target]) + ", " + kCType[ts[0]] + ">(" + A(0) + ")"));
  }
  if (op == "bitcast") {
    need(ts[0] != BOOL && target != BOOL && kSize[ts[0]] == kSize[target], "types of the same size");
    return scalar(target, bind(target, "ap::bitcast<" + std::string(kCType[target]) + ", " + kCType[ts[0]] + ">(" + A(0) + ")"));
  }
  err("unknown operator " + op);
}

std::string emit_function(const SX& fun, const std::vector<std::vector<int>>& ptypes,
                          const std::vector<std::vector<std::string>>& pexpr, const std::vector<int>& ret,
                          const std::string& ret_struct, Emitter em) {
  if (!fun.is("fun") || fun.l.size() != 3 || fun.l[1].atom) err("expected (fun (params...) body), got " + fun.str());
  const std::vector<SX>& ps = fun.l[1].l;
  if (ps.size() != ptypes.size())
    err("function of " + std::to_string(ps.size()) + " parameters where " + std::to_string(ptypes.size()) +
        " are expected: " + fun.str());
  for (size_t k = 0; k < ps.size(); ++k) {
    if (!ps[k].atom) err("parameter names must be symbols");
    Val v;
    v.t = ptypes[k];
    for (size_t f = 0; f < ptypes[k].size(); ++f)
      v.v.push_back(ptypes[k][f] == BOOL ? "(" + pexpr[k][f] + " != 0)" : pexpr[k][f]);
    em.env.emplace_back(ps[k].a, v);
  }
  const Val r = em.emit(fun.l[2]);
  if (r.t != ret) {
    std::string want, got;
    for (int t : ret) want += std::string(" ") + kName[t];
    for (int t : r.t) got += std::string(" ") + kName[t];
    err("function result has type (" + got + " ) where (" + want + " ) is expected: " + fun.str());
  }
  std::string s = em.body;
  if (ret_struct.empty()) {
    s += "    return " + r.v[0] + ";\n";
  } else {
    s += "    " + ret_struct + " r_;\n";
    for (size_t f = 0; f < ret.size(); ++f)
      s += "    r_.f" + std::to_string(f) + " = " +
           (ret[f] == BOOL ? "(uint8_t)(" + r.v[f] + " ? 1 : 0)" : "(" + std::string(kStore[ret[f]]) + ")" + r.v[f]) +
           ";\n";
    s += "    return r_;\n";
  }
  return s;
}

}  // namespace jit
}  // namespace acc
