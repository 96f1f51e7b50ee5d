// IR text format (see include/loop_tool/serialization.h).
//
//   var <name>
//   node <op> in=<i,..> vars=<v,..> prio=<p> order=<var:size:tail,..>
//        noreuse=<k,..> lannot=<s|s|..> annot=<s> syms=<slot=var,..>
//        cons=<expr>=<expr>;...
//   inputs <i,..>
//   outputs <o,..>
//
// Symbols inside constraints are numbered per node (slot #k); expressions
// are written in prefix form: (+ a b) (* a b) (/ a b) (% a b) (neg a)
// (rcp a) (max a b) (sz a), integers, and #k for symbol slot k.
// Annotation strings are percent-escaped.
#include "loop_tool/serialization.h"

#include <algorithm>
#include <cctype>
#include <sstream>

namespace loop_tool {

using symbolic::Expr;
using symbolic::Op;
using symbolic::Symbol;

namespace {

std::string escape(const std::string &s) {
  std::string out;
  for (unsigned char c : s) {
    if (std::isalnum(c) || c == '_' || c == '[' || c == ']' || c == '.' || c == '-') {
      out += (char)c;
    } else {
      char buf[4];
      snprintf(buf, sizeof(buf), "%%%02X", c);
      out += buf;
    }
  }
  return out;
}

std::string unescape(const std::string &s) {
  std::string out;
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] == '%' && i + 2 < s.size()) {
      out += (char)std::stoi(s.substr(i + 1, 2), nullptr, 16);
      i += 2;
    } else {
      out += s[i];
    }
  }
  return out;
}

std::vector<std::string> split(const std::string &s, char d) {
  std::vector<std::string> out;
  if (s.empty()) return out;
  std::string cur;
  for (char c : s) {
    if (c == d) {
      out.emplace_back(cur);
      cur.clear();
    } else {
      cur += c;
    }
  }
  out.emplace_back(cur);
  return out;
}

std::string write_expr(const Expr &e, std::unordered_map<int, int> &slots,
                       std::vector<Symbol> &slot_syms) {
  if (e.type() == Expr::Type::value) return std::to_string(e.value());
  if (e.type() == Expr::Type::symbol) {
    auto id = e.symbol().id();
    if (!slots.count(id)) {
      slots[id] = (int)slot_syms.size();
      slot_syms.emplace_back(e.symbol());
    }
    return "#" + std::to_string(slots.at(id));
  }
  const char *name = "";
  switch (e.op()) {
    case Op::add: name = "+"; break;
    case Op::multiply: name = "*"; break;
    case Op::divide: name = "/"; break;
    case Op::modulo: name = "%"; break;
    case Op::negate: name = "neg"; break;
    case Op::reciprocal: name = "rcp"; break;
    case Op::max: name = "max"; break;
    case Op::size: name = "sz"; break;
    default: ASSERT(0) << "cannot serialize " << e.dump();
  }
  std::string out = std::string("(") + name;
  for (const auto &a : e.args()) out += " " + write_expr(a, slots, slot_syms);
  return out + ")";
}

struct ExprReader {
  const std::string &s;
  size_t pos = 0;
  const std::vector<Symbol> &syms;
  void skip() {
    while (pos < s.size() && s[pos] == ' ') pos++;
  }
  Expr read() {
    skip();
    ASSERT(pos < s.size()) << "truncated expression";
    if (s[pos] == '#') {
      pos++;
      size_t n = 0;
      int k = std::stoi(s.substr(pos), &n);
      pos += n;
      return Expr(syms.at(k));
    }
    if (s[pos] != '(') {
      size_t n = 0;
      long long v = std::stoll(s.substr(pos), &n);
      pos += n;
      return Expr((int64_t)v);
    }
    pos++;
    size_t end = s.find(' ', pos);
    auto name = s.substr(pos, end - pos);
    pos = end;
    std::vector<Expr> args;
    while (true) {
      skip();
      if (s[pos] == ')') {
        pos++;
        break;
      }
      args.emplace_back(read());
    }
    if (name == "+") return Expr(Op::add, args.at(0), args.at(1));
    if (name == "*") return Expr(Op::multiply, args.at(0), args.at(1));
    if (name == "/") return Expr(Op::divide, args.at(0), args.at(1));
    if (name == "%") return Expr(Op::modulo, args.at(0), args.at(1));
    if (name == "max") return Expr(Op::max, args.at(0), args.at(1));
    if (name == "neg") return Expr(Op::negate, args.at(0));
    if (name == "rcp") return Expr(Op::reciprocal, args.at(0));
    if (name == "sz") return Expr(Op::size, args.at(0));
    ASSERT(0) << "unknown expression op " << name;
    return Expr(0);
  }
};

std::string join_ints(const std::vector<int> &v) {
  std::string out;
  for (size_t i = 0; i < v.size(); ++i) out += (i ? "," : "") + std::to_string(v[i]);
  return out;
}

std::vector<int> read_ints(const std::string &s) {
  std::vector<int> out;
  for (const auto &x : split(s, ',')) out.emplace_back(std::stoi(x));
  return out;
}

}  // namespace

std::string serialize(const IR &ir_in) {
  IR ir = ir_in;
  ir.reify_deletions();
  std::stringstream ss;
  for (auto v : ir.vars()) ss << "var " << escape(ir.var(v).name()) << "\n";
  for (auto nr : ir.nodes()) {
    const auto &n = ir.node(nr);
    ss << "node " << (int)n.op();
    ss << " in=" << join_ints(std::vector<int>(n.inputs().begin(), n.inputs().end()));
    ss << " vars=" << join_ints(std::vector<int>(n.vars().begin(), n.vars().end()));
    ss << " prio=" << ir.priority(nr);
    ss << " order=";
    const auto &order = ir.order(nr);
    for (size_t i = 0; i < order.size(); ++i)
      ss << (i ? "," : "") << order[i].first << ":" << order[i].second.size << ":"
         << order[i].second.tail;
    std::vector<int> noreuse(ir.not_reusable(nr).begin(), ir.not_reusable(nr).end());
    std::sort(noreuse.begin(), noreuse.end());
    ss << " noreuse=" << join_ints(noreuse);
    ss << " lannot=";
    auto la = ir.loop_annotations(nr);
    for (size_t i = 0; i < la.size(); ++i) ss << (i ? "|" : "") << escape(la[i]);
    ss << " annot=" << escape(ir.annotation(nr));
    std::unordered_map<int, int> slots;
    std::vector<Symbol> slot_syms;
    std::string cons;
    for (const auto &c : n.constraints())
      cons += write_expr(c.first, slots, slot_syms) + "=" +
              write_expr(c.second, slots, slot_syms) + ";";
    for (const auto &p : n.sym_to_var())
      if (!slots.count(p.first)) {
        slots[p.first] = (int)slot_syms.size();
        slot_syms.emplace_back(Symbol("s", p.first));
      }
    ss << " syms=";
    for (size_t k = 0; k < slot_syms.size(); ++k) {
      const auto &s = slot_syms[k];
      ss << (k ? "," : "") << escape(s.name()) << "/"
         << (n.has_sym(s) ? n.var(s) : -1);
    }
    ss << " cons=" << cons << "\n";
  }
  ss << "inputs " << join_ints(std::vector<int>(ir.inputs().begin(), ir.inputs().end()))
     << "\n";
  ss << "outputs "
     << join_ints(std::vector<int>(ir.outputs().begin(), ir.outputs().end())) << "\n";
  return ss.str();
}

IR deserialize(const std::string &str) {
  IR ir;
  struct Pending {
    int op;
    std::vector<int> inputs, vars;
    float prio;
    std::vector<std::pair<IR::VarRef, IR::LoopSize>> order;
    std::vector<int> noreuse;
    std::vector<std::string> lannot;
    std::string annot;
    std::vector<symbolic::Constraint> cons;
    std::unordered_map<int, IR::VarRef> sym_map;
  };
  std::vector<Pending> nodes;
  std::vector<int> inputs, outputs;
  std::stringstream in(str);
  std::string line;
  while (std::getline(in, line)) {
    if (line.rfind("var ", 0) == 0) {
      ir.create_var(unescape(line.substr(4)));
    } else if (line.rfind("node ", 0) == 0) {
      Pending p;
      std::stringstream ls(line.substr(5));
      ls >> p.op;
      std::string field;
      std::vector<Symbol> syms;
      std::string cons_text;
      while (ls >> field) {
        auto eq = field.find('=');
        auto key = field.substr(0, eq), val = field.substr(eq + 1);
        if (key == "in") p.inputs = read_ints(val);
        else if (key == "vars") p.vars = read_ints(val);
        else if (key == "prio") p.prio = std::stof(val);
        else if (key == "order") {
          for (const auto &o : split(val, ',')) {
            auto t = split(o, ':');
            p.order.emplace_back(std::stoi(t.at(0)),
                                 IR::LoopSize{std::stoll(t.at(1)), std::stoll(t.at(2))});
          }
        } else if (key == "noreuse") p.noreuse = read_ints(val);
        else if (key == "lannot") {
          for (const auto &a : split(val, '|')) p.lannot.emplace_back(unescape(a));
        } else if (key == "annot") p.annot = unescape(val);
        else if (key == "syms") {
          for (const auto &sv : split(val, ',')) {
            auto t = split(sv, '/');
            Symbol s(unescape(t.at(0)));
            syms.emplace_back(s);
            int v = std::stoi(t.at(1));
            if (v >= 0) p.sym_map[s.id()] = v;
          }
        } else if (key == "cons") {
          cons_text = val;
          // the rest of the line belongs to the constraint list
          std::string rest;
          std::getline(ls, rest);
          cons_text += rest;
        }
      }
      for (const auto &c : split(cons_text, ';')) {
        if (c.empty()) continue;
        // split at the top-level '='
        int depth = 0;
        size_t at = std::string::npos;
        for (size_t i = 0; i < c.size(); ++i) {
          if (c[i] == '(') depth++;
          if (c[i] == ')') depth--;
          if (c[i] == '=' && depth == 0) {
            at = i;
            break;
          }
        }
        ASSERT(at != std::string::npos) << "bad constraint text " << c;
        auto l = c.substr(0, at), r = c.substr(at + 1);
        ExprReader lr{l, 0, syms}, rr{r, 0, syms};
        p.cons.emplace_back(lr.read(), rr.read());
      }
      if (p.lannot.size() != p.order.size()) p.lannot.assign(p.order.size(), "");
      nodes.emplace_back(std::move(p));
    } else if (line.rfind("inputs", 0) == 0) {
      inputs = line.size() > 7 ? read_ints(line.substr(7)) : std::vector<int>{};
    } else if (line.rfind("outputs", 0) == 0) {
      outputs = line.size() > 8 ? read_ints(line.substr(8)) : std::vector<int>{};
    }
  }
  for (auto &p : nodes) {
    std::vector<IR::NodeRef> ins(p.inputs.begin(), p.inputs.end());
    std::vector<IR::VarRef> vars(p.vars.begin(), p.vars.end());
    auto nr = ir.create_node((Operation)p.op, ins, vars, p.cons, p.sym_map);
    ir.set_priority(nr, p.prio);
    ir.set_order(nr, p.order, p.lannot);
    for (auto k : p.noreuse) ir.disable_reuse(nr, k);
    ir.annotate(nr, p.annot);
  }
  ir.set_inputs(std::vector<IR::NodeRef>(inputs.begin(), inputs.end()));
  ir.set_outputs(std::vector<IR::NodeRef>(outputs.begin(), outputs.end()));
  return ir;
}

}  // namespace loop_tool
