// libzkp_fused_codegen.so -- ArithGraph (serde JSON) -> HIP fused elementwise kernels.
//
// Counterpart of the reference compiler's FusedOp::{new, gen} (core/src/fused_kernels/
// kernel_gen.rs:47-618), which turns one common::arith::ArithGraph
// (common/src/arith.rs:172-425) into per-element CUDA code.  Here the graph
// arrives as its serde JSON (externally tagged enums, as serde_json writes the
// Rust types) and the output is HIP for gfx950:
//
//   extern "C" hipError_t <name>(const ConstPolyPtr* vars, const PolyPtr* mut_vars,
//                                unsigned long long len, bool is_first,
//                                void* local_buffer, hipStream_t stream);
//
// with the reference's argument convention (kernel_gen.rs:286-290,
// core/src/fused_kernels.rs:40-52): vars = the graph's inputs in order minus
// in-place inputs, mut_vars = its outputs in order, an in-place (Mut) input read
// through the mut_vars slot of the output that overwrites it (kernel_gen.rs:61-83,
// common/src/arith.rs:411-452).  Scalars are 1-element views.  Semantics per
// operation follow kernel_gen.rs:457-566, including the coefficient-form rule
// for scalar-polynomial add/sub (only element 0 of the first chunk sees the
// scalar).  Register allocation is left to the compiler: the straight-line body
// needs no explicit spill buffer, so local_buffer is accepted and unused.
//
// Each library also exports <name>_batched, which runs the same expression over
// several independent instances of different lengths in one launch.
#include <errno.h>
#include <fcntl.h>
#include <spawn.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/wait.h>

#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <string>
#include <vector>

extern char** environ;

namespace {

// ------------------------------------------------------------------ JSON

struct Json {
  enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
  bool b = false;
  double num = 0;
  uint64_t u = 0;
  bool is_uint = false;
  std::string s;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;

  const Json* get(const std::string& k) const {
    for (auto& kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct Parser {
  const char* p;
  const char* end;
  std::string err;

  void ws() {
    while (p < end && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++;
  }
  bool fail(const std::string& m) {
    if (err.empty()) err = m;
    return false;
  }
  bool parse(Json& out) {
    ws();
    if (p >= end) return fail("unexpected end of JSON");
    char c = *p;
    if (c == '{') {
      out.kind = Json::Obj;
      p++;
      ws();
      if (p < end && *p == '}') {
        p++;
        return true;
      }
      while (true) {
        ws();
        Json key;
        if (p >= end || *p != '"' || !parse(key)) return fail("expected object key");
        ws();
        if (p >= end || *p != ':') return fail("expected ':'");
        p++;
        Json val;
        if (!parse(val)) return false;
        out.obj.emplace_back(key.s, std::move(val));
        ws();
        if (p < end && *p == ',') {
          p++;
          continue;
        }
        if (p < end && *p == '}') {
          p++;
          return true;
        }
        return fail("expected ',' or '}'");
      }
    }
    if (c == '[') {
      out.kind = Json::Arr;
      p++;
      ws();
      if (p < end && *p == ']') {
        p++;
        return true;
      }
      while (true) {
        Json val;
        if (!parse(val)) return false;
        out.arr.push_back(std::move(val));
        ws();
        if (p < end && *p == ',') {
          p++;
          continue;
        }
        if (p < end && *p == ']') {
          p++;
          return true;
        }
        return fail("expected ',' or ']'");
      }
    }
    if (c == '"') {
      out.kind = Json::Str;
      p++;
      while (p < end && *p != '"') {
        if (*p == '\\' && p + 1 < end) {
          p++;
          char e = *p;
          out.s.push_back(e == 'n' ? '\n' : e == 't' ? '\t' : e);
        } else {
          out.s.push_back(*p);
        }
        p++;
      }
      if (p >= end) return fail("unterminated string");
      p++;
      return true;
    }
    if (!strncmp(p, "null", 4)) {
      p += 4;
      out.kind = Json::Null;
      return true;
    }
    if (!strncmp(p, "true", 4)) {
      p += 4;
      out.kind = Json::Bool;
      out.b = true;
      return true;
    }
    if (!strncmp(p, "false", 5)) {
      p += 5;
      out.kind = Json::Bool;
      return true;
    }
    if (c == '-' || (c >= '0' && c <= '9')) {
      const char* s = p;
      bool neg = c == '-';
      if (neg) p++;
      uint64_t v = 0;
      bool integral = true;
      while (p < end && *p >= '0' && *p <= '9') v = v * 10 + (uint64_t)(*p++ - '0');
      while (p < end && (*p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-' || (*p >= '0' && *p <= '9'))) {
        integral = false;
        p++;
      }
      out.kind = Json::Num;
      out.num = strtod(std::string(s, p).c_str(), nullptr);
      out.u = v;
      out.is_uint = integral && !neg;
      return true;
    }
    return fail(std::string("unexpected character '") + c + "'");
  }
};

// ------------------------------------------------------------------ graph

enum class OpKind { Input, Output, Bin, Unr, Todo };
enum class Group { Pp, Ss, Sp, P, S };

struct Vertex {
  OpKind kind = OpKind::Todo;
  // Input / Output
  bool scalar = false;  // FusedType::Scalar
  bool mut_input = false;
  uint64_t store_node = 0;
  bool has_in_node = false;
  uint64_t in_node = 0;
  // Arith
  Group group = Group::Pp;
  std::string op;  // Add Sub SubBy Mul Div DivBy Neg Inv Pow
  uint64_t pow = 0;
  uint64_t lhs = 0, rhs = 0;
};

struct Graph {
  std::vector<uint64_t> inputs, outputs;
  std::vector<Vertex> v;
  bool coef = false;
};

bool as_index(const Json& j, uint64_t* out) {
  // a usize id is a number; accept a one-element array too (newtype written as tuple)
  if (j.kind == Json::Num && j.is_uint) {
    *out = j.u;
    return true;
  }
  if (j.kind == Json::Arr && j.arr.size() == 1) return as_index(j.arr[0], out);
  return false;
}

bool parse_graph(const Json& root, Graph& g, std::string& err) {
  if (root.kind != Json::Obj) {
    err = "graph JSON must be an object";
    return false;
  }
  const Json* outs = root.get("outputs");
  const Json* ins = root.get("inputs");
  const Json* gg = root.get("g");
  const Json* repr = root.get("poly_repr");
  if (!outs || !ins || !gg || !repr) {
    err = "graph JSON needs outputs, inputs, g and poly_repr";
    return false;
  }
  for (auto& x : outs->arr) {
    uint64_t id;
    if (!as_index(x, &id)) return (err = "bad output id", false);
    g.outputs.push_back(id);
  }
  for (auto& x : ins->arr) {
    uint64_t id;
    if (!as_index(x, &id)) return (err = "bad input id", false);
    g.inputs.push_back(id);
  }
  g.coef = repr->kind == Json::Str && repr->s == "Coef";
  if (repr->kind != Json::Str || (repr->s != "Coef" && repr->s != "Lagrange"))
    return (err = "poly_repr must be Coef or Lagrange", false);
  // Digraph(Heap(Vec<V>, PhantomData)) serialises as [[...], null]; accept [...] too
  const Json* verts = gg;
  if (gg->kind == Json::Arr && gg->arr.size() == 2 && gg->arr[0].kind == Json::Arr && gg->arr[1].kind == Json::Null)
    verts = &gg->arr[0];
  if (verts->kind != Json::Arr) return (err = "g must be a vertex list", false);
  for (auto& vj : verts->arr) {
    const Json* op = vj.kind == Json::Obj && vj.get("op") ? vj.get("op") : &vj;
    Vertex v;
    if (op->kind == Json::Str && op->s == "Todo") {
      v.kind = OpKind::Todo;
    } else if (op->kind == Json::Obj && op->obj.size() == 1) {
      const std::string& tag = op->obj[0].first;
      const Json& body = op->obj[0].second;
      if (tag == "Input") {
        v.kind = OpKind::Input;
        const Json* typ = body.get("typ");
        const Json* mut = body.get("mutability");
        if (!typ || !mut) return (err = "Input needs typ and mutability", false);
        v.scalar = typ->s == "Scalar";
        v.mut_input = mut->s == "Mut";
      } else if (tag == "Output") {
        v.kind = OpKind::Output;
        const Json* typ = body.get("typ");
        const Json* sn = body.get("store_node");
        const Json* in = body.get("in_node");
        if (!typ || !sn || !as_index(*sn, &v.store_node)) return (err = "Output needs typ and store_node", false);
        v.scalar = typ->s == "Scalar";
        if (in && in->kind != Json::Null) {
          if (!as_index(*in, &v.in_node)) return (err = "bad in_node", false);
          v.has_in_node = true;
        }
      } else if (tag == "Arith") {
        if (body.kind != Json::Obj || body.obj.size() != 1) return (err = "bad Arith", false);
        const std::string& at = body.obj[0].first;
        const Json& args = body.obj[0].second;
        if (at == "Bin") {
          v.kind = OpKind::Bin;
          if (args.kind != Json::Arr || args.arr.size() != 3) return (err = "Bin needs [op, lhs, rhs]", false);
          const Json& bop = args.arr[0];
          if (bop.kind != Json::Obj || bop.obj.size() != 1) return (err = "bad BinOp", false);
          const std::string& grp = bop.obj[0].first;
          v.group = grp == "Pp" ? Group::Pp : grp == "Ss" ? Group::Ss : Group::Sp;
          if (grp != "Pp" && grp != "Ss" && grp != "Sp") return (err = "BinOp group must be Pp/Ss/Sp", false);
          v.op = bop.obj[0].second.s;
          if (!as_index(args.arr[1], &v.lhs) || !as_index(args.arr[2], &v.rhs)) return (err = "bad Bin operands", false);
        } else if (at == "Unr") {
          v.kind = OpKind::Unr;
          if (args.kind != Json::Arr || args.arr.size() != 2) return (err = "Unr needs [op, arg]", false);
          const Json& uop = args.arr[0];
          if (uop.kind != Json::Obj || uop.obj.size() != 1) return (err = "bad UnrOp", false);
          const std::string& grp = uop.obj[0].first;
          if (grp != "P" && grp != "S") return (err = "UnrOp group must be P/S", false);
          v.group = grp == "P" ? Group::P : Group::S;
          const Json& o = uop.obj[0].second;
          if (o.kind == Json::Str) {
            v.op = o.s;
          } else if (o.kind == Json::Obj && o.obj.size() == 1 && o.obj[0].first == "Pow") {
            v.op = "Pow";
            if (!as_index(o.obj[0].second, &v.pow)) return (err = "bad Pow exponent", false);
          } else {
            return (err = "bad unary op", false);
          }
          if (!as_index(args.arr[1], &v.lhs)) return (err = "bad Unr operand", false);
        } else {
          return (err = "Arith must be Bin or Unr", false);
        }
      } else {
        return (err = "unknown operation " + tag, false);
      }
    } else {
      return (err = "bad vertex", false);
    }
    g.v.push_back(v);
  }
  return true;
}

// ------------------------------------------------------------------ emitter

struct Emitter {
  const Graph& g;
  std::string name;
  std::string err;
  std::map<uint64_t, size_t> var_slot;      // input id -> vars index
  std::map<uint64_t, size_t> mut_slot;      // input/output id -> mut_vars index
  std::set<size_t> read_slots;              // mut_vars the kernel also reads (in-place inputs)
  std::vector<uint64_t> order;              // topological order of used vertices
  std::set<uint64_t> done;

  explicit Emitter(const Graph& gr, const std::string& n) : g(gr), name(n) {}

  bool valid(uint64_t id) { return id < g.v.size(); }

  bool visit(uint64_t id, std::set<uint64_t>& stack) {
    if (!valid(id)) return (err = "vertex id out of range", false);
    if (done.count(id)) return true;
    if (stack.count(id)) return (err = "graph has a cycle", false);
    stack.insert(id);
    const Vertex& x = g.v[id];
    if (x.kind == OpKind::Bin) {
      if (!visit(x.lhs, stack) || !visit(x.rhs, stack)) return false;
    } else if (x.kind == OpKind::Unr) {
      if (!visit(x.lhs, stack)) return false;
    } else if (x.kind == OpKind::Output) {
      if (!visit(x.store_node, stack)) return false;
    } else if (x.kind == OpKind::Todo) {
      return (err = "Todo vertex reached by an output", false);
    }
    stack.erase(id);
    done.insert(id);
    order.push_back(id);
    return true;
  }

  bool plan() {
    // var lists (common/src/arith.rs:428-452 gen_var_lists)
    std::set<uint64_t> inplace;
    for (size_t k = 0; k < g.outputs.size(); k++) {
      uint64_t o = g.outputs[k];
      if (!valid(o) || g.v[o].kind != OpKind::Output) return (err = "outputs must be Output vertices", false);
      mut_slot[o] = k;
      const Vertex& ov = g.v[o];
      if (ov.has_in_node) {
        if (!valid(ov.in_node) || g.v[ov.in_node].kind != OpKind::Input) return (err = "in_node must be an Input", false);
        if (g.v[ov.in_node].mut_input) {
          inplace.insert(ov.in_node);
          mut_slot[ov.in_node] = k;
          read_slots.insert(k);
        }
      }
    }
    size_t slot = 0;
    for (uint64_t i : g.inputs) {
      if (!valid(i) || g.v[i].kind != OpKind::Input) return (err = "inputs must be Input vertices", false);
      if (inplace.count(i)) continue;
      var_slot[i] = slot++;
    }
    std::set<uint64_t> stack;
    for (uint64_t o : g.outputs)
      if (!visit(o, stack)) return false;
    // type checks and unsupported ops (kernel_gen.rs:480-566)
    for (uint64_t id : order) {
      const Vertex& x = g.v[id];
      if (x.kind == OpKind::Bin && (x.op == "Div" || x.op == "DivBy"))
        return (err = "division is not fused (handled by batched/scalar invert)", false);
      if (x.kind == OpKind::Unr && x.op == "Inv") return (err = "inversion is not fused (batched invert)", false);
      if (x.kind == OpKind::Unr && x.group == Group::S && x.op == "Pow")
        return (err = "scalar power is not fused (scalar pow)", false);
      if (x.kind == OpKind::Bin && x.op != "Add" && x.op != "Sub" && x.op != "SubBy" && x.op != "Mul")
        return (err = "unknown binary op " + x.op, false);
      if (x.kind == OpKind::Bin && x.op == "SubBy" && x.group != Group::Sp)
        return (err = "SubBy only exists for scalar-polynomial ops", false);
      if (x.kind == OpKind::Input && !var_slot.count(id) && !mut_slot.count(id))
        return (err = "input vertex not listed in inputs", false);
    }
    return true;
  }

  static std::string tmp(uint64_t id) { return "t" + std::to_string(id); }

  // is the value of vertex id per element (polynomial) rather than one scalar?
  bool is_poly(uint64_t id) const {
    const Vertex& x = g.v[id];
    switch (x.kind) {
      case OpKind::Input:
        return !x.scalar;
      case OpKind::Bin:
        return x.group != Group::Ss;
      case OpKind::Unr:
        return x.group == Group::P;
      default:
        return false;
    }
  }

  std::string expr(uint64_t id) const {
    const Vertex& x = g.v[id];
    const std::string l = tmp(x.lhs), r = x.kind == OpKind::Bin ? tmp(x.rhs) : "";
    if (x.kind == OpKind::Unr) {
      if (x.op == "Neg") return l + ".neg()";
      return l + ".pow(" + std::to_string(x.pow) + "ull)";
    }
    const char* sym = x.op == "Add" ? " + " : x.op == "Mul" ? " * " : " - ";
    if (x.group != Group::Sp) return l + sym + r;
    // scalar (lhs) with polynomial (rhs)
    if (x.op == "Mul") return l + " * " + r;
    if (!g.coef) {
      if (x.op == "Add") return l + " + " + r;
      if (x.op == "Sub") return l + " - " + r;
      return r + " - " + l;  // SubBy
    }
    const std::string head = "(idx == 0 && is_first)";
    if (x.op == "Add") return head + " ? " + l + " + " + r + " : " + r;
    if (x.op == "Sub") return head + " ? " + l + " - " + r + " : " + r + ".neg()";
    return head + " ? " + r + " - " + l + " : " + r;  // SubBy
  }

  std::string src_ptr(uint64_t id) const {
    auto it = var_slot.find(id);
    if (it != var_slot.end()) return "vars[" + std::to_string(it->second) + "]";
    return "mut_vars[" + std::to_string(mut_slot.at(id)) + "]";
  }

  // body shared by the single and batched kernels
  void emit_body(std::ostringstream& o, const std::string& vars, const std::string& mut_vars) const {
    // views and loop-invariant scalars
    for (uint64_t id : order) {
      const Vertex& x = g.v[id];
      if (x.kind == OpKind::Input) {
        std::string p = src_ptr(id);
        p = (p[0] == 'v' ? vars : mut_vars) + p.substr(p.find('['));
        if (x.scalar)
          o << "  const zkp::Fr " << tmp(id) << " = zkp::Fr::load(" << p << ".ptr);\n";
        else
          o << "  const zkp::PolyView v" << id << " = zkp::fused::view(" << p << ");\n";
      } else if (x.kind == OpKind::Output) {
        const std::string p = mut_vars + "[" + std::to_string(mut_slot.at(id)) + "]";
        if (x.scalar)
          o << "  uint32_t* const o" << id << " = " << p << ".ptr;\n";
        else
          o << "  const zkp::PolyView o" << id << " = zkp::fused::view(" << p << ");\n";
      } else if (!is_poly(id)) {
        o << "  const zkp::Fr " << tmp(id) << " = " << expr(id) << ";\n";
      }
    }
  }

  void emit_element(std::ostringstream& o, const char* indent) const {
    for (uint64_t id : order) {
      const Vertex& x = g.v[id];
      if (x.kind == OpKind::Input && !x.scalar) {
        o << indent << "const zkp::Fr " << tmp(id) << " = zkp::Fr::load(v" << id << ".elem(idx));\n";
      } else if ((x.kind == OpKind::Bin || x.kind == OpKind::Unr) && is_poly(id)) {
        o << indent << "const zkp::Fr " << tmp(id) << " = " << expr(id) << ";\n";
      }
    }
    // stores after all loads: an in-place output never clobbers an input still to be read
    for (uint64_t id : order) {
      const Vertex& x = g.v[id];
      if (x.kind != OpKind::Output) continue;
      if (x.scalar)
        o << indent << "if (idx == 0) " << tmp(x.store_node) << ".store(o" << id << ");\n";
      else
        o << indent << tmp(x.store_node) << ".store(o" << id << ".elem(idx));\n";
    }
  }

  std::string source() const {
    std::ostringstream o;
    o << "// Generated fused elementwise kernel " << name << " (" << order.size() << " vertices, "
      << (g.coef ? "coefficient" : "Lagrange") << " form).\n";
    o << "#include \"fused/fused_runtime.hpp\"\n\n";
    o << "namespace {\n";
    o << "__global__ void __launch_bounds__(zkp::fused::THREADS)\n    " << name
      << "_kernel(const ConstPolyPtr* __restrict__ vars, const PolyPtr* __restrict__ mut_vars, "
         "unsigned long long len, bool is_first) {\n";
    emit_body(o, "vars", "mut_vars");
    o << "  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;\n";
    o << "  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < len; idx += stride) {\n";
    emit_element(o, "    ");
    o << "  }\n}\n\n";
    // batched: instance k uses vars[k * NV ..], mut_vars[k * NM ..] and elements [prefix[k], prefix[k+1])
    const size_t nv = var_slot.size(), nm = g.outputs.size();
    o << "__global__ void __launch_bounds__(zkp::fused::THREADS)\n    " << name
      << "_batched_kernel(const ConstPolyPtr* __restrict__ vars_all, const PolyPtr* __restrict__ mut_vars_all, "
         "const unsigned long long* __restrict__ prefix, unsigned int instances, bool is_first) {\n";
    o << "  const uint64_t total = prefix[instances];\n";
    o << "  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;\n";
    o << "  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {\n";
    o << "    const unsigned int k = zkp::fused::instance_of(prefix, instances, g);\n";
    o << "    const uint64_t idx = g - prefix[k];\n";
    o << "    const ConstPolyPtr* vars = vars_all + (uint64_t)k * " << nv << ";\n";
    o << "    const PolyPtr* mut_vars = mut_vars_all + (uint64_t)k * " << nm << ";\n";
    o << "    (void)vars;\n";
    std::ostringstream body;
    emit_body(body, "vars", "mut_vars");
    std::string b = body.str();
    // indent the body by two more spaces
    std::string ind;
    for (size_t i = 0; i < b.size(); i++) {
      if (i == 0 || b[i - 1] == '\n') ind += "  ";
      ind += b[i];
    }
    o << ind;
    emit_element(o, "    ");
    o << "  }\n}\n}  // namespace\n\n";
    o << "extern \"C\" hipError_t " << name
      << "(const ConstPolyPtr* vars, const PolyPtr* mut_vars, unsigned long long len, bool is_first, "
         "void* local_buffer, hipStream_t stream) {\n";
    o << "  (void)local_buffer;\n";
    o << "  if (len == 0) return hipSuccess;\n";
    o << "  if (mut_vars == nullptr" << (nv ? " || vars == nullptr" : "") << ") return hipErrorInvalidValue;\n";
    o << "  hipLaunchKernelGGL(" << name
      << "_kernel, dim3(zkp::fused::grid_for(len)), dim3(zkp::fused::THREADS), 0, stream, vars, mut_vars, len, "
         "is_first);\n";
    o << "  return hipGetLastError();\n}\n\n";
    o << "// instances k < n: vars[k*" << nv << " ..], mut_vars[k*" << nm
      << " ..] (device arrays), prefix[k] = sum of earlier lengths (device, n+1 entries),\n"
         "// total = prefix[n] (host copy, sizes the grid).\n";
    o << "extern \"C\" hipError_t " << name
      << "_batched(const ConstPolyPtr* vars, const PolyPtr* mut_vars, const unsigned long long* prefix, "
         "unsigned int instances, unsigned long long total, bool is_first, hipStream_t stream) {\n";
    o << "  if (instances == 0 || total == 0) return hipSuccess;\n";
    o << "  if (mut_vars == nullptr || prefix == nullptr" << (nv ? " || vars == nullptr" : "")
      << ") return hipErrorInvalidValue;\n";
    o << "  hipLaunchKernelGGL(" << name
      << "_batched_kernel, dim3(zkp::fused::grid_for(total)), dim3(zkp::fused::THREADS), 0, stream, vars, "
         "mut_vars, prefix, instances, is_first);\n";
    o << "  return hipGetLastError();\n}\n\n";
    o << "// number of vars / mut_vars the kernel expects\n";
    o << "extern \"C\" void " << name << "_arity(unsigned int* out2) { out2[0] = " << nv << "; out2[1] = " << nm
      << "; }\n";
    // which mut_vars hold data the kernel reads (in-place inputs); the others
    // are written in full, so a streaming caller need not load them
    o << "extern \"C\" void " << name << "_mut_reads(unsigned char* out) {";
    for (size_t k = 0; k < nm; k++) o << " out[" << k << "] = " << (read_slots.count(k) ? 1 : 0) << ";";
    o << " (void)out; }\n";
    return o.str();
  }
};

bool valid_name(const char* n) {
  if (!n || !*n) return false;
  if (!((n[0] >= 'a' && n[0] <= 'z') || (n[0] >= 'A' && n[0] <= 'Z') || n[0] == '_')) return false;
  for (const char* p = n; *p; p++)
    if (!((*p >= 'a' && *p <= 'z') || (*p >= 'A' && *p <= 'Z') || (*p >= '0' && *p <= '9') || *p == '_'))
      return false;
  return true;
}

void set_err(char* err, size_t cap, const std::string& m) {
  if (err && cap) {
    strncpy(err, m.c_str(), cap - 1);
    err[cap - 1] = 0;
  }
}

bool generate(const char* graph_json, const char* name, std::string& out, std::string& err) {
  if (!graph_json) return (err = "null graph", false);
  if (!valid_name(name)) return (err = "kernel name must be a C identifier", false);
  Parser ps{graph_json, graph_json + strlen(graph_json), ""};
  Json root;
  if (!ps.parse(root)) return (err = "JSON: " + ps.err, false);
  Graph g;
  if (!parse_graph(root, g, err)) return false;
  Emitter e(g, name);
  if (!e.plan()) return (err = e.err, false);
  out = e.source();
  return true;
}

// Runs argv with stdout and stderr sent to log_path; on failure `log` gets the
// tail of that file.
int run_process(const std::vector<std::string>& argv, const std::string& log_path, std::string& log) {
  std::vector<char*> args;
  for (auto& a : argv) args.push_back(const_cast<char*>(a.c_str()));
  args.push_back(nullptr);
  posix_spawn_file_actions_t fa;
  posix_spawn_file_actions_init(&fa);
  posix_spawn_file_actions_addopen(&fa, 1, log_path.c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
  posix_spawn_file_actions_adddup2(&fa, 1, 2);
  pid_t pid;
  int rc = posix_spawnp(&pid, args[0], &fa, nullptr, args.data(), environ);
  posix_spawn_file_actions_destroy(&fa);
  if (rc != 0) {
    log = std::string("cannot start ") + argv[0] + ": " + strerror(rc);
    return -1;
  }
  int status = 0;
  while (waitpid(pid, &status, 0) < 0 && errno == EINTR) {
  }
  const int code = WIFEXITED(status) ? WEXITSTATUS(status) : -1;
  if (code != 0) {
    FILE* f = fopen(log_path.c_str(), "r");
    if (f) {
      std::string all;
      char buf[4096];
      size_t n;
      while ((n = fread(buf, 1, sizeof buf, f)) > 0) all.append(buf, n);
      fclose(f);
      log = all.size() > 3000 ? all.substr(all.size() - 3000) : all;
    }
  }
  return code;
}

}  // namespace

extern "C" {

// HIP source of one fused kernel; 0 on success, -1 with a message in err.
// out_len receives the full length (copy truncated to out_cap - 1 bytes).
int zkp_fused_codegen(const char* graph_json, const char* name, char* out, size_t out_cap, size_t* out_len,
                      char* err, size_t err_cap) {
  std::string src, e;
  if (!generate(graph_json, name, src, e)) {
    set_err(err, err_cap, e);
    return -1;
  }
  if (out_len) *out_len = src.size();
  if (out && out_cap) {
    size_t n = src.size() < out_cap - 1 ? src.size() : out_cap - 1;
    memcpy(out, src.data(), n);
    out[n] = 0;
  }
  return 0;
}

// Generates every kernel into out_dir/<name>.hip and compiles them with hipcc
// into out_dir/libfused_kernels.so (the per-program library the reference's
// FusedKernel loads, core/src/fused_kernels.rs:38,76-95).  include_dir is the
// csrc directory holding fused/fused_runtime.hpp; extra_flags (may be NULL) is
// appended to the hipcc line.  0 on success, -1 with a message in err.
int zkp_fused_build(const char* const* graph_jsons, const char* const* names, int count, const char* out_dir,
                    const char* hipcc, const char* include_dir, const char* const* extra_flags, int n_extra,
                    char* err, size_t err_cap) {
  if (count <= 0 || !out_dir || !hipcc || !include_dir) {
    set_err(err, err_cap, "bad arguments");
    return -1;
  }
  mkdir(out_dir, 0755);
  std::vector<std::string> argv = {hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
                                   "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                                   "-Wno-unused-command-line-argument", std::string("-I") + include_dir};
  for (int i = 0; i < n_extra; i++) argv.push_back(extra_flags[i]);
  std::set<std::string> seen;
  for (int k = 0; k < count; k++) {
    std::string src, e;
    if (!generate(graph_jsons[k], names[k], src, e)) {
      set_err(err, err_cap, std::string(names[k] ? names[k] : "?") + ": " + e);
      return -1;
    }
    if (!seen.insert(names[k]).second) {
      set_err(err, err_cap, std::string("duplicate kernel name ") + names[k]);
      return -1;
    }
    const std::string path = std::string(out_dir) + "/" + names[k] + ".hip";
    FILE* f = fopen(path.c_str(), "w");
    if (!f) {
      set_err(err, err_cap, "cannot write " + path);
      return -1;
    }
    fwrite(src.data(), 1, src.size(), f);
    fclose(f);
    argv.push_back(path);
  }
  const std::string lib = std::string(out_dir) + "/libfused_kernels.so";
  const std::string tmp = lib + ".tmp";
  argv.push_back("-o");
  argv.push_back(tmp);
  std::string log;
  const int rc = run_process(argv, std::string(out_dir) + "/build.log", log);
  if (rc != 0) {
    set_err(err, err_cap, log.empty() ? "hipcc failed with status " + std::to_string(rc) : log);
    return -1;
  }
  if (rename(tmp.c_str(), lib.c_str()) != 0) {
    set_err(err, err_cap, "cannot rename " + tmp);
    return -1;
  }
  return 0;
}

}  // extern "C"
