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
#include <limits.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/wait.h>
#include <unistd.h>

#include <algorithm>
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

  Emitter(const Graph& gr, const std::string& n, size_t budget, unsigned waves)
      : g(gr), name(n), reg_budget(budget < 3 ? 3 : budget), waves_per_eu(waves) {}

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
  bool is_scalar_value(uint64_t id) const {
    const Vertex& x = g.v[id];
    return (x.kind == OpKind::Input || x.kind == OpKind::Bin || x.kind == OpKind::Unr) && !is_poly(id);
  }

  // the expression of an arithmetic vertex over operand names l, r
  std::string expr(uint64_t id, const std::string& l, const std::string& r) const {
    const Vertex& x = g.v[id];
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

  std::string src_ptr(uint64_t id, const std::string& vars, const std::string& mut_vars) const {
    auto it = var_slot.find(id);
    if (it != var_slot.end()) return vars + "[" + std::to_string(it->second) + "]";
    return mut_vars + "[" + std::to_string(mut_slot.at(id)) + "]";
  }

  // ---------------------------------------------------------------- scalars
  //
  // Values that are one field element for the whole launch (scalar inputs and
  // Ss / scalar-Neg results).  With few of them they are computed before the
  // element loop and held in registers.  Many of them (a compiler's fused
  // graph carries dozens) would stay live across the loop and crowd out the
  // per-element values, so above SCALAR_REGS the single-instance entry point
  // computes them once in a one-thread prologue kernel into a table at the
  // head of local_buffer, which the element loop reads at each use (uniform
  // addresses: scalar loads); scalar outputs are stored by that prologue,
  // after every scalar input has been read.
  static constexpr size_t SCALAR_REGS = 4;
  std::vector<uint64_t> scalar_order;       // scalar vertices in evaluation order
  std::map<uint64_t, size_t> scalar_index;  // vertex -> table slot

  void plan_scalars() {
    for (uint64_t id : order)
      if (is_scalar_value(id)) {
        scalar_index[id] = scalar_order.size();
        scalar_order.push_back(id);
      }
  }
  bool scalar_table() const { return scalar_order.size() > SCALAR_REGS; }
  size_t table_bytes() const { return scalar_table() ? align256(scalar_order.size() * 32) : 0; }
  static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

  // straight-line evaluation of every scalar vertex into const locals t<id>
  void emit_scalar_values(std::ostringstream& o, const std::string& vars, const std::string& mut_vars,
                          const char* ind) const {
    for (uint64_t id : scalar_order) {
      const Vertex& x = g.v[id];
      if (x.kind == OpKind::Input)
        o << ind << "const zkp::Fr " << tmp(id) << " = zkp::Fr::load(" << src_ptr(id, vars, mut_vars) << ".ptr);\n";
      else
        o << ind << "const zkp::Fr " << tmp(id) << " = "
          << expr(id, tmp(x.lhs), x.kind == OpKind::Bin ? tmp(x.rhs) : "") << ";\n";
    }
  }

  // ---------------------------------------------------------------- per-element schedule and registers
  //
  // The element loop evaluates the per-element vertices in the depth-first
  // order of the outputs, holding at most reg_budget field elements in
  // registers (8 VGPRs each; a product needs about 40 more), so that the
  // kernel keeps several waves per SIMD.  When a value must leave, the one used
  // furthest in the future goes (Belady): an input polynomial is simply read
  // again later, a computed value is written once to its spill slot in
  // local_buffer (two 16-byte planes per slot, lane-contiguous, so spills are
  // coalesced) and read back before its next use.  Outputs are stored as soon
  // as their value exists, except an output that overwrites an input still to
  // be read, which waits for the end of the element.
  size_t reg_budget = 16;     // field elements kept in registers per element
  unsigned waves_per_eu = 2;  // occupancy the register allocator must keep (0: compiler's choice)
  // grid when spilling: the workgroups of 256 threads that fit 256 CUs at the
  // kernel's occupancy (four per CU when the compiler chooses)
  unsigned spill_blocks() const { return 256u * (waves_per_eu ? waves_per_eu : 4u); }
  std::vector<uint64_t> elem_order;  // per-element vertices and poly outputs, in order
  size_t spill_slots = 0;
  std::string element_code;  // generated with `indent` placeholders resolved later

  struct Alloc {
    const Emitter& e;
    std::ostringstream& o;
    const std::string ind;
    std::map<uint64_t, std::vector<size_t>> uses;  // vertex -> positions of its uses
    std::map<uint64_t, std::string> reg;           // vertex -> name of its value in a register
    std::map<uint64_t, size_t> slot;               // spilled vertex -> slot
    std::vector<size_t> free_slots;
    size_t slots = 0;
    size_t counter = 0;
    size_t pos = 0;

    size_t next_use(uint64_t id) const {
      auto it = uses.find(id);
      if (it == uses.end()) return SIZE_MAX;
      for (size_t u : it->second)
        if (u > pos) return u;
      return SIZE_MAX;
    }
    bool used_at_or_after(uint64_t id, size_t t) const {
      auto it = uses.find(id);
      if (it == uses.end()) return false;
      for (size_t u : it->second)
        if (u >= t) return true;
      return false;
    }
    std::string fresh(uint64_t id) { return "t" + std::to_string(id) + "_" + std::to_string(counter++); }

    void evict_one(const std::set<uint64_t>& keep) {
      uint64_t victim = UINT64_MAX;
      size_t far = 0;
      for (auto& kv : reg) {
        if (keep.count(kv.first)) continue;
        const size_t nu = next_use(kv.first);
        if (victim == UINT64_MAX || nu > far) {
          victim = kv.first;
          far = nu;
        }
      }
      if (victim == UINT64_MAX) return;
      const Vertex& x = e.g.v[victim];
      if (far != SIZE_MAX && x.kind != OpKind::Input && !slot.count(victim)) {
        size_t sl;
        if (!free_slots.empty()) {
          sl = free_slots.back();
          free_slots.pop_back();
        } else {
          sl = slots++;
        }
        slot[victim] = sl;
        o << ind << "zkp::fused::spill_store(spill, " << sl << ", kSpillThreads, " << reg[victim] << ");\n";
      }
      reg.erase(victim);
    }
    void make_room(const std::set<uint64_t>& keep) {
      while (reg.size() >= e.reg_budget) {
        const size_t before = reg.size();
        evict_one(keep);
        if (reg.size() == before) break;  // everything is pinned by the current operation
      }
    }
    // name of a register holding the value of per-element vertex id
    std::string get(uint64_t id, const std::set<uint64_t>& keep) {
      auto it = reg.find(id);
      if (it != reg.end()) return it->second;
      make_room(keep);
      const std::string n = fresh(id);
      const Vertex& x = e.g.v[id];
      if (x.kind == OpKind::Input)
        o << ind << "const zkp::Fr " << n << " = zkp::Fr::load(v" << id << ".elem(idx));\n";
      else
        o << ind << "const zkp::Fr " << n << " = zkp::fused::spill_load(spill, " << slot.at(id) << ", kSpillThreads);\n";
      reg[id] = n;
      return n;
    }
    // values with no use after the current position leave registers (and slots)
    void retire(uint64_t id) {
      if (next_use(id) != SIZE_MAX) return;
      reg.erase(id);
      auto it = slot.find(id);
      if (it != slot.end()) {
        free_slots.push_back(it->second);
        slot.erase(it);
      }
    }
  };

  std::string operand(const Alloc& A, uint64_t id, const std::map<uint64_t, std::string>& names) const {
    (void)A;
    if (is_scalar_value(id)) {
      if (scalar_table()) return "zkp::Fr::load(scalars + " + std::to_string(scalar_index.at(id) * 8) + ")";
      return tmp(id);
    }
    return names.at(id);
  }

  void plan_elements() {
    std::set<uint64_t> inplace_inputs;  // mutable inputs some output overwrites
    for (uint64_t id : order) {
      const Vertex& x = g.v[id];
      if ((x.kind == OpKind::Bin || x.kind == OpKind::Unr) && is_poly(id)) elem_order.push_back(id);
      if (x.kind == OpKind::Output && !x.scalar) elem_order.push_back(id);
    }
  }

  // Element code (one pass of the allocator); `ind` is the indentation.
  std::string gen_element(const char* ind_c) {
    std::ostringstream o;
    Alloc A{*this, o, std::string(ind_c), {}, {}, {}, {}, 0, 0, 0};
    // positions of uses; deferred in-place outputs use their value at the end
    const size_t END = elem_order.size();
    std::vector<uint64_t> deferred;
    std::map<uint64_t, size_t> last_read;  // input -> last position reading it
    for (size_t t = 0; t < elem_order.size(); t++) {
      const Vertex& x = g.v[elem_order[t]];
      auto note = [&](uint64_t c) {
        if (!is_poly(c)) return;
        A.uses[c].push_back(t);
        if (g.v[c].kind == OpKind::Input) last_read[c] = t;
      };
      if (x.kind == OpKind::Bin) {
        note(x.lhs);
        note(x.rhs);
      } else if (x.kind == OpKind::Unr) {
        note(x.lhs);
      }
    }
    for (size_t t = 0; t < elem_order.size(); t++) {
      const uint64_t id = elem_order[t];
      const Vertex& x = g.v[id];
      if (x.kind != OpKind::Output) continue;
      bool defer = false;
      if (x.has_in_node && g.v[x.in_node].mut_input) {
        auto it = last_read.find(x.in_node);
        defer = it != last_read.end() && it->second > t;
      }
      if (defer) {
        deferred.push_back(id);
        A.uses[x.store_node].push_back(END);
      } else if (is_poly(x.store_node)) {
        A.uses[x.store_node].push_back(t);
      }
    }
    for (auto& kv : A.uses) std::sort(kv.second.begin(), kv.second.end());
    auto store = [&](uint64_t out_id, const std::set<uint64_t>& keep) {
      const Vertex& x = g.v[out_id];
      std::string val;
      if (is_poly(x.store_node)) val = A.get(x.store_node, keep);
      else val = operand(A, x.store_node, {});
      o << ind_c << val << ".store(o" << out_id << ".elem(idx));\n";
    };
    std::set<uint64_t> deferred_set(deferred.begin(), deferred.end());
    for (size_t t = 0; t < elem_order.size(); t++) {
      A.pos = t;
      const uint64_t id = elem_order[t];
      const Vertex& x = g.v[id];
      if (x.kind == OpKind::Output) {
        if (!deferred_set.count(id)) {
          store(id, {x.store_node});
          if (is_poly(x.store_node)) A.retire(x.store_node);
        }
        continue;
      }
      std::set<uint64_t> keep;
      if (is_poly(x.lhs)) keep.insert(x.lhs);
      if (x.kind == OpKind::Bin && is_poly(x.rhs)) keep.insert(x.rhs);
      std::map<uint64_t, std::string> names;
      for (uint64_t c : keep) names[c] = A.get(c, keep);
      const std::string l = operand(A, x.lhs, names);
      const std::string r = x.kind == OpKind::Bin ? operand(A, x.rhs, names) : "";
      for (uint64_t c : keep) A.retire(c);
      A.make_room(keep);
      const std::string n = A.fresh(id);
      o << ind_c << "const zkp::Fr " << n << " = " << expr(id, l, r) << ";\n";
      A.reg[id] = n;
      A.retire(id);  // never used (an output of nothing): drop at once
    }
    A.pos = END;
    for (uint64_t d : deferred) store(d, {g.v[d].store_node});
    spill_slots = A.slots > spill_slots ? A.slots : spill_slots;
    return o.str();
  }

  // views (and, without a scalar table, the scalar values) before the element loop
  void emit_views(std::ostringstream& o, const std::string& vars, const std::string& mut_vars, const char* ind,
                  bool scalars_inline) const {
    for (uint64_t id : order) {
      const Vertex& x = g.v[id];
      if (x.kind == OpKind::Input && !x.scalar)
        o << ind << "const zkp::PolyView v" << id << " = zkp::fused::view(" << src_ptr(id, vars, mut_vars) << ");\n";
      else if (x.kind == OpKind::Output && !x.scalar)
        o << ind << "const zkp::PolyView o" << id << " = zkp::fused::view(" << mut_vars << "["
          << mut_slot.at(id) << "]);\n";
    }
    if (scalars_inline) emit_scalar_values(o, vars, mut_vars, ind);
  }

  // scalar outputs: in the loop at element 0 (inline scalars) or in the prologue
  void emit_scalar_outputs(std::ostringstream& o, const std::string& mut_vars, const char* ind,
                           const char* guard) const {
    for (uint64_t id : order) {
      const Vertex& x = g.v[id];
      if (x.kind != OpKind::Output || !x.scalar) continue;
      o << ind << guard << tmp(x.store_node) << ".store(" << mut_vars << "[" << mut_slot.at(id) << "].ptr);\n";
    }
  }

  std::string source() {
    plan_scalars();
    plan_elements();
    const bool table = scalar_table();
    const std::string elem_single = gen_element("    ");
    const std::string elem_batched = gen_element("    ");  // same allocation, own names
    const bool spills = spill_slots > 0;
    const size_t nv = var_slot.size(), nm = g.outputs.size();
    const unsigned blocks = spill_blocks();
    const size_t local_bytes = table_bytes() + (spills ? spill_slots * 32ull * blocks * 256ull : 0);
    std::ostringstream o;
    o << "// Generated fused elementwise kernel " << name << " (" << order.size() << " vertices, "
      << (g.coef ? "coefficient" : "Lagrange") << " form; " << scalar_order.size() << " scalar values"
      << (table ? " in a table" : "") << ", " << spill_slots << " spill slots).\n";
    o << "#include \"fused/fused_runtime.hpp\"\n\n";
    o << "namespace {\n";
    if (spills) o << "constexpr uint64_t kSpillThreads = " << blocks << "ull * zkp::fused::THREADS;\n\n";
    if (table) {
      o << "__global__ void " << name
        << "_scalars_kernel(const ConstPolyPtr* __restrict__ vars, const PolyPtr* __restrict__ mut_vars, "
           "uint32_t* __restrict__ table) {\n";
      emit_scalar_values(o, "vars", "mut_vars", "  ");
      for (uint64_t id : scalar_order) o << "  " << tmp(id) << ".store(table + " << scalar_index.at(id) * 8 << ");\n";
      emit_scalar_outputs(o, "mut_vars", "  ", "");
      o << "}\n\n";
    }
    // the occupancy the register budget was chosen for binds only kernels that spill; a small kernel keeps
    // whatever occupancy its few registers give (its loads need the waves)
    const std::string bounds = "__launch_bounds__(zkp::fused::THREADS)" +
        (waves_per_eu && spills ? " __attribute__((amdgpu_waves_per_eu(" + std::to_string(waves_per_eu) + ")))"
                                : std::string());
    o << "__global__ void " << bounds << "\n    " << name
      << "_kernel(const ConstPolyPtr* __restrict__ vars, const PolyPtr* __restrict__ mut_vars, "
         "unsigned long long len, bool is_first, uint32_t* __restrict__ local) {\n";
    o << "  (void)local;\n";
    if (table) o << "  const uint32_t* __restrict__ scalars = local;\n";
    if (spills)
      o << "  uint4* __restrict__ spill = reinterpret_cast<uint4*>(reinterpret_cast<char*>(local) + " << table_bytes()
        << ") + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;\n";
    emit_views(o, "vars", "mut_vars", "  ", !table);
    o << "  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;\n";
    o << "  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < len; idx += stride) {\n";
    o << elem_single;
    if (!table) emit_scalar_outputs(o, "mut_vars", "    ", "if (idx == 0) ");
    o << "  }\n}\n\n";
    // batched: instance k uses vars[k * NV ..], mut_vars[k * NM ..] and elements [prefix[k], prefix[k+1]);
    // its scalars are evaluated per element
    o << "__global__ void " << bounds << "\n    " << name
      << "_batched_kernel(const ConstPolyPtr* __restrict__ vars_all, const PolyPtr* __restrict__ mut_vars_all, "
         "const unsigned long long* __restrict__ prefix, unsigned int instances, bool is_first, "
         "uint32_t* __restrict__ local) {\n";
    o << "  (void)local;\n";
    if (spills)
      o << "  uint4* __restrict__ spill = reinterpret_cast<uint4*>(reinterpret_cast<char*>(local) + " << table_bytes()
        << ") + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;\n";
    o << "  const uint64_t total = prefix[instances];\n";
    o << "  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;\n";
    o << "  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {\n";
    o << "    const unsigned int k = zkp::fused::instance_of(prefix, instances, g);\n";
    o << "    const uint64_t idx = g - prefix[k];\n";
    o << "    const ConstPolyPtr* vars = vars_all + (uint64_t)k * " << nv << ";\n";
    o << "    const PolyPtr* mut_vars = mut_vars_all + (uint64_t)k * " << nm << ";\n";
    o << "    (void)vars;\n";
    emit_views(o, "vars", "mut_vars", "    ", true);
    if (table) {
      // no table per instance: the element code reads scalars from one; give it a local copy
      o << "    uint32_t scalars_buf[" << scalar_order.size() * 8 << "];\n";
      for (uint64_t id : scalar_order) o << "    " << tmp(id) << ".store(scalars_buf + " << scalar_index.at(id) * 8 << ");\n";
      o << "    const uint32_t* scalars = scalars_buf;\n";
    }
    o << elem_batched;
    emit_scalar_outputs(o, "mut_vars", "    ", "if (idx == 0) ");
    o << "  }\n}\n}  // namespace\n\n";
    const char* grid = spills ? "zkp::fused::grid_capped(len, " : "zkp::fused::grid_for(len";
    o << "extern \"C\" hipError_t " << name
      << "(const ConstPolyPtr* vars, const PolyPtr* mut_vars, unsigned long long len, bool is_first, "
         "void* local_buffer, hipStream_t stream) {\n";
    o << "  if (len == 0) return hipSuccess;\n";
    o << "  if (mut_vars == nullptr" << (nv ? " || vars == nullptr" : "") << (local_bytes ? " || local_buffer == nullptr" : "")
      << ") return hipErrorInvalidValue;\n";
    if (table)
      o << "  hipLaunchKernelGGL(" << name << "_scalars_kernel, dim3(1), dim3(1), 0, stream, vars, mut_vars, "
           "static_cast<uint32_t*>(local_buffer));\n";
    o << "  hipLaunchKernelGGL(" << name << "_kernel, dim3(" << grid << (spills ? std::to_string(blocks) : "")
      << ")), dim3(zkp::fused::THREADS), 0, stream, vars, mut_vars, len, is_first, "
         "static_cast<uint32_t*>(local_buffer));\n";
    o << "  return hipGetLastError();\n}\n\n";
    o << "// instances k < n: vars[k*" << nv << " ..], mut_vars[k*" << nm
      << " ..] (device arrays), prefix[k] = sum of earlier lengths (device, n+1 entries),\n"
         "// total = prefix[n] (host copy, sizes the grid); local_buffer as for the single entry point.\n";
    o << "extern \"C\" hipError_t " << name
      << "_batched(const ConstPolyPtr* vars, const PolyPtr* mut_vars, const unsigned long long* prefix, "
         "unsigned int instances, unsigned long long total, bool is_first, hipStream_t stream, void* local_buffer) {\n";
    o << "  if (instances == 0 || total == 0) return hipSuccess;\n";
    o << "  if (mut_vars == nullptr || prefix == nullptr" << (nv ? " || vars == nullptr" : "")
      << (spills ? " || local_buffer == nullptr" : "") << ") return hipErrorInvalidValue;\n";
    o << "  hipLaunchKernelGGL(" << name << "_batched_kernel, dim3("
      << (spills ? std::string("zkp::fused::grid_capped(total, ") + std::to_string(blocks) + ")"
                 : std::string("zkp::fused::grid_for(total)"))
      << "), dim3(zkp::fused::THREADS), 0, stream, vars, mut_vars, prefix, instances, is_first, "
         "static_cast<uint32_t*>(local_buffer));\n";
    o << "  return hipGetLastError();\n}\n\n";
    o << "// number of vars / mut_vars the kernel expects\n";
    o << "extern \"C\" void " << name << "_arity(unsigned int* out2) { out2[0] = " << nv << "; out2[1] = " << nm
      << "; }\n";
    // which mut_vars hold data the kernel reads (in-place inputs); the others
    // are written in full, so a streaming caller need not load them
    o << "extern \"C\" void " << name << "_mut_reads(unsigned char* out) {";
    for (size_t k = 0; k < nm; k++) o << " out[" << k << "] = " << (read_slots.count(k) ? 1 : 0) << ";";
    o << " (void)out; }\n";
    o << "// bytes of local_buffer the entry points need (0: it may be NULL): the scalar table, then the\n"
         "// spill slots of every thread of the capped grid\n";
    o << "extern \"C\" unsigned long long " << name << "_local_buffer_bytes(void) { return " << local_bytes
      << "ull; }\n";
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

// code generation options (zkp_fused_set_options)
size_t g_reg_budget = 16;
unsigned g_waves_per_eu = 2;

bool generate(const char* graph_json, const char* name, std::string& out, std::string& err) {
  if (!graph_json) return (err = "null graph", false);
  if (!valid_name(name)) return (err = "kernel name must be a C identifier", false);
  Parser ps{graph_json, graph_json + strlen(graph_json), ""};
  Json root;
  if (!ps.parse(root)) return (err = "JSON: " + ps.err, false);
  Graph g;
  if (!parse_graph(root, g, err)) return false;
  Emitter e(g, name, g_reg_budget, g_waves_per_eu);
  if (!e.plan()) return (err = e.err, false);
  out = e.source();
  return true;
}

// Starts argv with stdout and stderr sent to log_path; the pid, or -1 (with `log` set).
pid_t spawn_process(const std::vector<std::string>& argv, const std::string& log_path, std::string& log) {
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
  return pid;
}

// Exit status of pid; on failure `log` gets the tail of log_path.
int wait_process(pid_t pid, const std::string& log_path, std::string& log) {
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

int run_process(const std::vector<std::string>& argv, const std::string& log_path, std::string& log) {
  const pid_t pid = spawn_process(argv, log_path, log);
  return pid < 0 ? -1 : wait_process(pid, log_path, log);
}

}  // namespace

extern "C" {

// Options of the kernels generated afterwards (process-wide, for the code
// generator only): reg_budget field elements held in registers per element
// before values spill to local_buffer (at least 3), and the waves per SIMD the
// compiler must keep (0: its own choice).
void zkp_fused_set_options(unsigned int reg_budget, unsigned int waves_per_eu) {
  g_reg_budget = reg_budget;
  g_waves_per_eu = waves_per_eu;
}

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
  std::vector<std::string> base = {hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
                                   "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                                   "-Wno-unused-command-line-argument", std::string("-I") + include_dir};
  for (int i = 0; i < n_extra; i++) base.push_back(extra_flags[i]);
  // every kernel compiled on its own, up to `jobs` at a time (large graphs take minutes each), then one link
  std::set<std::string> seen;
  std::vector<std::string> objects;
  std::vector<std::pair<pid_t, std::string>> running;  // pid, log path
  const long cpus = sysconf(_SC_NPROCESSORS_ONLN);
  const size_t jobs = cpus > 1 ? (size_t)(cpus > 16 ? 16 : cpus) : 1;
  std::string first_err;
  auto reap_one = [&]() {
    const auto job = running.front();
    running.erase(running.begin());
    std::string log;
    if (wait_process(job.first, job.second, log) != 0 && first_err.empty())
      first_err = log.empty() ? "hipcc failed (" + job.second + ")" : log;
  };
  for (int k = 0; k < count; k++) {
    std::string src, e;
    if (!generate(graph_jsons[k], names[k], src, e)) {
      set_err(err, err_cap, std::string(names[k] ? names[k] : "?") + ": " + e);
      while (!running.empty()) reap_one();
      return -1;
    }
    if (!seen.insert(names[k]).second) {
      set_err(err, err_cap, std::string("duplicate kernel name ") + names[k]);
      while (!running.empty()) reap_one();
      return -1;
    }
    const std::string path = std::string(out_dir) + "/" + names[k] + ".hip";
    FILE* f = fopen(path.c_str(), "w");
    if (!f) {
      set_err(err, err_cap, "cannot write " + path);
      while (!running.empty()) reap_one();
      return -1;
    }
    fwrite(src.data(), 1, src.size(), f);
    fclose(f);
    const std::string obj = std::string(out_dir) + "/" + names[k] + ".o";
    std::vector<std::string> argv = base;
    argv.insert(argv.end(), {"-c", path, "-o", obj});
    if (running.size() >= jobs) reap_one();
    std::string log;
    const std::string log_path = std::string(out_dir) + "/" + names[k] + ".log";
    const pid_t pid = spawn_process(argv, log_path, log);
    if (pid < 0) {
      if (first_err.empty()) first_err = log;
    } else {
      running.emplace_back(pid, log_path);
    }
    objects.push_back(obj);
  }
  while (!running.empty()) reap_one();
  if (!first_err.empty()) {
    set_err(err, err_cap, first_err);
    return -1;
  }
  const std::string lib = std::string(out_dir) + "/libfused_kernels.so";
  const std::string tmp = lib + ".tmp";
  std::vector<std::string> link = {hipcc, "--offload-arch=gfx950", "-shared", "-fPIC"};
  link.insert(link.end(), objects.begin(), objects.end());
  link.push_back("-o");
  link.push_back(tmp);
  std::string log;
  const int rc = run_process(link, std::string(out_dir) + "/build.log", log);
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
