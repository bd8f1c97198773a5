// The program's own functions, as text, to HIP C++: the scalar expression
// language of the run-time compiled skeletons.  A host prints a function of
// the Accelerate AST (`Fun`/`OpenExp`, whose primitives are the PrimFun
// constructors dispatched at LLVM/CodeGen/Exp.hs:244-310) as an
// S-expression; this translation checks its types and emits straight-line
// device code calling the semantics of jit/prelude.hpp.
#pragma once
#include <stdint.h>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace acc {
namespace jit {

// Scalar types (Accelerate SingleType; Int = i64, Word = u64; Bool is stored
// as one byte, PrimBool, LLVM/CodeGen/Exp.hs:187-193).
enum Sc : int { I8, I16, I32, I64, U8, U16, U32, U64, F32, F64, BOOL, NSC };

const char* sc_name(int t);
const char* sc_ctype(int t);      // type of a value in generated code
const char* sc_storetype(int t);  // type of an array field in memory
int sc_size(int t);
int sc_parse(const std::string& s);  // -1 if not a type name
inline bool sc_int(int t) { return t >= I8 && t <= U64; }
inline bool sc_signed(int t) { return t >= I8 && t <= I64; }
inline bool sc_float(int t) { return t == F32 || t == F64; }
inline bool sc_num(int t) { return t >= I8 && t <= F64; }

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// S-expressions
struct SX {
  bool atom = false;
  bool quoted = false;  // a string literal ("..."): `a` holds its text
  std::string a;
  std::vector<SX> l;
  bool is(const char* head) const { return !atom && !l.empty() && l[0].atom && l[0].a == head; }
  std::string str() const;
};
SX parse_sx(const std::string& text);

// A value in generated code: one C expression per tuple field.
struct Val {
  std::vector<std::string> v;
  std::vector<int> t;
};

// Translates expressions to statements appended to `body`.
struct Emitter {
  std::string body;
  int tmp = 0;
  std::vector<std::pair<std::string, Val>> env;
  std::vector<int> free_types;  // element type of free array k (`(read k i)`)
  // stencil functions: (nb dy dx) reads the neighbourhood around `nb_ptr`
  bool stencil = false;
  int nb_ry = 0, nb_rx = 0, nb_stride = 0, nb_type = F32;
  // foreign scalar functions met so far: name -> file-scope definition
  // (shared by every copy of the prototype emitter of one node)
  std::shared_ptr<std::map<std::string, std::string>> foreign;

  Val emit(const SX& e);
  std::string bind(int t, const std::string& expr);
};

// Byte layout of a struct of the given scalar fields (C rules), and of the
// same struct followed by a 32-bit flag (Maybe).
struct Layout {
  int size = 0, align = 1, maybe_size = 0;
};
Layout layout_of(const std::vector<int>& fields);

// `struct NAME { T0 f0; ... };` with storage types
std::string struct_decl(const char* name, const std::vector<int>& fields);

// A device function `RET NAME(PARAMS) const` body from (fun (p...) e): params
// bound to C expressions `pexpr` with element types `ptypes`; the result must
// have type `ret` and is returned as struct `ret_struct` (or a scalar when
// ret_struct is empty).
std::string emit_function(const SX& fun, const std::vector<std::vector<int>>& ptypes,
                          const std::vector<std::vector<std::string>>& pexpr, const std::vector<int>& ret,
                          const std::string& ret_struct, Emitter proto);

}  // namespace jit
}  // namespace acc
