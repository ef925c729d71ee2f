// Configuration files: parseConfig / loadConfig (src/io/Config.cpp:83-293 of
// the reference, which reads TOML with toml++). toml++ is not available here,
// so this file carries its own reader for the part of TOML a run
// configuration uses: [table] and [[array-of-tables]] headers, bare or quoted
// keys, basic and literal strings, integers, floats, booleans, arrays
// (spanning lines, trailing comma allowed) and inline tables. The mapping onto
// RunConfig follows the reference key for key: the same defaults, the same
// scalar-or-per-spin field rule, step fields falling back to the baseline,
// and a ConfigError naming the key for every problem. [ensemble] also takes
// min_separation and secular_angle, and [[coupling]] tables give fixed
// couplings: options RunConfig carries beyond the reference's.
#include <cctype>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <iterator>
#include <memory>
#include <optional>
#include <string>
#include <utility>
#include <vector>

#include "spinsim/io/Config.hpp"

namespace spinsim {
namespace {

// ------------------------------------------------------------- the values

struct TomlValue;
using TomlArray = std::vector<TomlValue>;

struct TomlTable {
  std::vector<std::pair<std::string, std::unique_ptr<TomlValue>>> entries;

  [[nodiscard]] const TomlValue* find(std::string_view key) const {
    for (const auto& [k, v] : entries)
      if (k == key) return v.get();
    return nullptr;
  }
  [[nodiscard]] bool contains(std::string_view key) const { return find(key) != nullptr; }
};

struct TomlValue {
  enum class Kind { Integer, Float, Boolean, String, Array, Table } kind = Kind::Integer;
  std::int64_t integer = 0;
  double real = 0.0;
  bool boolean = false;
  std::string text;
  TomlArray array;
  std::unique_ptr<TomlTable> table;
  bool header_table = false;   // defined by a [header] (may not be redefined)

  // Conversions a lookup allows: integers read as floats, floats that hold
  // an integer exactly read as integers; nothing else converts.
  [[nodiscard]] std::optional<double> as_double() const {
    if (kind == Kind::Float) return real;
    if (kind == Kind::Integer) return static_cast<double>(integer);
    return std::nullopt;
  }
  [[nodiscard]] std::optional<std::int64_t> as_int() const {
    if (kind == Kind::Integer) return integer;
    if (kind == Kind::Float && std::isfinite(real) && std::trunc(real) == real && std::abs(real) < 9.2e18)
      return static_cast<std::int64_t>(real);
    return std::nullopt;
  }
  [[nodiscard]] std::optional<bool> as_bool() const {
    if (kind == Kind::Boolean) return boolean;
    return std::nullopt;
  }
  [[nodiscard]] std::optional<std::string> as_string() const {
    if (kind == Kind::String) return text;
    return std::nullopt;
  }
};

TomlValue make_table() {
  TomlValue v;
  v.kind = TomlValue::Kind::Table;
  v.table = std::make_unique<TomlTable>();
  return v;
}

// ------------------------------------------------------------- the reader

class Reader {
 public:
  Reader(std::string_view text, std::string_view origin) : s_(text), origin_(origin) {}

  TomlValue document() {
    TomlValue root = make_table();
    TomlTable* current = root.table.get();
    for (;;) {
      skip_blank_lines();
      if (at_end()) break;
      if (peek() == '[') {
        const bool array_of_tables = peek(1) == '[';
        pos_ += array_of_tables ? 2 : 1;
        skip_space();
        const std::string name = key();
        skip_space();
        expect(']');
        if (array_of_tables) expect(']');
        end_of_line();
        current = open_header(*root.table, name, array_of_tables);
      } else {
        key_value(*current);
        end_of_line();
      }
    }
    return root;
  }

 private:
  [[noreturn]] void error(const std::string& what) const {
    throw ConfigError(origin_, "<syntax>", "line " + std::to_string(line_) + ": " + what);
  }
  [[nodiscard]] bool at_end() const { return pos_ >= s_.size(); }
  [[nodiscard]] char peek(std::size_t ahead = 0) const {
    return pos_ + ahead < s_.size() ? s_[pos_ + ahead] : '\0';
  }
  void expect(char c) {
    if (peek() != c) error(std::string("expected '") + c + "'");
    ++pos_;
  }
  void skip_space() {
    while (peek() == ' ' || peek() == '\t') ++pos_;
  }
  void skip_comment() {
    if (peek() == '#')
      while (!at_end() && peek() != '\n') ++pos_;
  }
  // Blank lines, comments, and (inside arrays) line breaks.
  void skip_blank_lines() {
    for (;;) {
      skip_space();
      skip_comment();
      if (peek() == '\r' && peek(1) == '\n') ++pos_;
      if (peek() == '\n') {
        ++pos_;
        ++line_;
        continue;
      }
      return;
    }
  }
  void end_of_line() {
    skip_space();
    skip_comment();
    if (peek() == '\r') ++pos_;
    if (at_end()) return;
    if (peek() != '\n') error("unexpected text after a value");
    ++pos_;
    ++line_;
  }

  static bool bare_key_char(char c) {
    return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') || c == '_' || c == '-';
  }
  std::string key() {
    if (peek() == '"') return basic_string();
    if (peek() == '\'') return literal_string();
    const std::size_t b = pos_;
    while (bare_key_char(peek())) ++pos_;
    if (pos_ == b) error("expected a key");
    if (peek() == '.') error("dotted keys are not supported");
    return std::string(s_.substr(b, pos_ - b));
  }

  TomlTable* open_header(TomlTable& root, const std::string& name, bool array_of_tables) {
    for (auto& [k, v] : root.entries) {
      if (k != name) continue;
      if (array_of_tables) {
        if (v->kind != TomlValue::Kind::Array) error("'" + name + "' is already defined and is not an array of tables");
        v->array.push_back(make_table());
        return v->array.back().table.get();
      }
      if (v->kind != TomlValue::Kind::Table || v->header_table) error("table '" + name + "' is defined twice");
      v->header_table = true;
      return v->table.get();
    }
    auto v = std::make_unique<TomlValue>();
    if (array_of_tables) {
      v->kind = TomlValue::Kind::Array;
      v->array.push_back(make_table());
      TomlTable* t = v->array.back().table.get();
      root.entries.emplace_back(name, std::move(v));
      return t;
    }
    *v = make_table();
    v->header_table = true;
    TomlTable* t = v->table.get();
    root.entries.emplace_back(name, std::move(v));
    return t;
  }

  void key_value(TomlTable& into) {
    const std::string k = key();
    skip_space();
    expect('=');
    skip_space();
    if (into.contains(k)) error("key '" + k + "' is defined twice");
    auto v = std::make_unique<TomlValue>(value());
    into.entries.emplace_back(k, std::move(v));
  }

  TomlValue value() {
    const char c = peek();
    if (c == '"' || c == '\'') {
      TomlValue v;
      v.kind = TomlValue::Kind::String;
      v.text = c == '"' ? basic_string() : literal_string();
      return v;
    }
    if (c == '[') return array();
    if (c == '{') return inline_table();
    if (s_.substr(pos_, 4) == "true" && !bare_key_char(peek(4))) {
      pos_ += 4;
      TomlValue v;
      v.kind = TomlValue::Kind::Boolean;
      v.boolean = true;
      return v;
    }
    if (s_.substr(pos_, 5) == "false" && !bare_key_char(peek(5))) {
      pos_ += 5;
      TomlValue v;
      v.kind = TomlValue::Kind::Boolean;
      return v;
    }
    return number();
  }

  std::string basic_string() {
    expect('"');
    if (peek() == '"' && peek(1) == '"') error("multi-line strings are not supported");
    std::string out;
    for (;;) {
      if (at_end() || peek() == '\n') error("unterminated string");
      const char c = s_[pos_++];
      if (c == '"') return out;
      if (c != '\\') {
        out.push_back(c);
        continue;
      }
      const char e = at_end() ? '\0' : s_[pos_++];
      switch (e) {
        case 'b': out.push_back('\b'); break;
        case 't': out.push_back('\t'); break;
        case 'n': out.push_back('\n'); break;
        case 'f': out.push_back('\f'); break;
        case 'r': out.push_back('\r'); break;
        case '"': out.push_back('"'); break;
        case '\\': out.push_back('\\'); break;
        default: error("unsupported escape sequence");
      }
    }
  }
  std::string literal_string() {
    expect('\'');
    const std::size_t b = pos_;
    while (!at_end() && peek() != '\'' && peek() != '\n') ++pos_;
    if (peek() != '\'') error("unterminated string");
    std::string out(s_.substr(b, pos_ - b));
    ++pos_;
    return out;
  }

  TomlValue number() {
    const std::size_t b = pos_;
    while (!at_end()) {
      const char c = peek();
      if ((c >= '0' && c <= '9') || c == '+' || c == '-' || c == '.' || c == '_' || c == 'e' || c == 'E' ||
          c == 'i' || c == 'n' || c == 'f' || c == 'a')
        ++pos_;
      else
        break;
    }
    std::string tok(s_.substr(b, pos_ - b));
    if (tok.empty()) error("expected a value");
    // underscores only between digits
    std::string digits;
    for (std::size_t i = 0; i < tok.size(); ++i) {
      if (tok[i] != '_') {
        digits.push_back(tok[i]);
        continue;
      }
      const bool ok = i > 0 && i + 1 < tok.size() && std::isdigit(static_cast<unsigned char>(tok[i - 1])) &&
                      std::isdigit(static_cast<unsigned char>(tok[i + 1]));
      if (!ok) error("misplaced '_' in '" + tok + "'");
    }
    TomlValue v;
    const std::string body = (digits[0] == '+' || digits[0] == '-') ? digits.substr(1) : digits;
    if (body == "inf" || body == "nan") {
      v.kind = TomlValue::Kind::Float;
      v.real = body == "inf" ? INFINITY : NAN;
      if (digits[0] == '-') v.real = -v.real;
      return v;
    }
    const bool is_int = !body.empty() && body.find_first_not_of("0123456789") == std::string::npos;
    if (is_int) {
      if (body.size() > 1 && body[0] == '0') error("leading zeros in '" + tok + "'");
      try {
        std::size_t used = 0;
        v.integer = std::stoll(digits, &used, 10);
        if (used != digits.size()) error("malformed integer '" + tok + "'");
      } catch (const std::out_of_range&) {
        error("integer '" + tok + "' out of range");
      }
      v.kind = TomlValue::Kind::Integer;
      return v;
    }
    // a float: digits, then a fraction and/or an exponent
    const std::size_t dot = body.find('.'), ex = body.find_first_of("eE");
    const bool digit_first = !body.empty() && std::isdigit(static_cast<unsigned char>(body[0]));
    const bool fraction_ok = dot == std::string::npos ||
                             (dot + 1 < body.size() && std::isdigit(static_cast<unsigned char>(body[dot + 1])));
    if (!digit_first || !fraction_ok || (dot == std::string::npos && ex == std::string::npos))
      error("malformed number '" + tok + "'");
    try {
      std::size_t used = 0;
      v.real = std::stod(digits, &used);
      if (used != digits.size()) error("malformed number '" + tok + "'");
    } catch (const std::invalid_argument&) {
      error("malformed number '" + tok + "'");
    } catch (const std::out_of_range&) {
      error("number '" + tok + "' out of range");
    }
    v.kind = TomlValue::Kind::Float;
    return v;
  }

  TomlValue array() {
    expect('[');
    TomlValue v;
    v.kind = TomlValue::Kind::Array;
    for (;;) {
      skip_blank_lines();
      if (peek() == ']') {
        ++pos_;
        return v;
      }
      if (at_end()) error("unterminated array");
      v.array.push_back(value());
      skip_blank_lines();
      if (peek() == ',') {
        ++pos_;
        continue;
      }
      if (peek() == ']') {
        ++pos_;
        return v;
      }
      error("expected ',' or ']' in an array");
    }
  }

  TomlValue inline_table() {
    expect('{');
    TomlValue v = make_table();
    skip_space();
    if (peek() == '}') {
      ++pos_;
      return v;
    }
    for (;;) {
      skip_space();
      key_value(*v.table);
      skip_space();
      if (peek() == ',') {
        ++pos_;
        continue;
      }
      if (peek() == '}') {
        ++pos_;
        return v;
      }
      error("expected ',' or '}' in an inline table");
    }
  }

  std::string_view s_;
  std::string_view origin_;
  std::size_t pos_ = 0;
  std::size_t line_ = 1;
};

// ---------------------------------------------- the mapping onto RunConfig

const TomlValue* lookup(const TomlTable* t, std::string_view key) { return t ? t->find(key) : nullptr; }

const TomlTable* section(const TomlTable& root, std::string_view name) {
  const TomlValue* v = root.find(name);
  return v && v->kind == TomlValue::Kind::Table ? v->table.get() : nullptr;
}

const TomlArray* array_at(const TomlTable* t, std::string_view key) {
  const TomlValue* v = lookup(t, key);
  return v && v->kind == TomlValue::Kind::Array ? &v->array : nullptr;
}

std::string require_string(const TomlTable* t, std::string_view key, std::string_view origin) {
  const TomlValue* v = lookup(t, key);
  const std::optional<std::string> s = v ? v->as_string() : std::nullopt;
  if (!s) throw ConfigError(origin, key, "missing, or not of the expected type");
  return *s;
}
double require_double(const TomlTable* t, std::string_view key, std::string_view origin) {
  const TomlValue* v = lookup(t, key);
  const std::optional<double> d = v ? v->as_double() : std::nullopt;
  if (!d) throw ConfigError(origin, key, "missing, or not of the expected type");
  return *d;
}
std::int64_t require_int(const TomlTable* t, std::string_view key, std::string_view origin) {
  const TomlValue* v = lookup(t, key);
  const std::optional<std::int64_t> i = v ? v->as_int() : std::nullopt;
  if (!i) throw ConfigError(origin, key, "missing, or not of the expected type");
  return *i;
}
// A present value of the wrong type falls back too, as in the reference.
std::string optional_string(const TomlTable* t, std::string_view key, std::string fallback) {
  const TomlValue* v = lookup(t, key);
  const std::optional<std::string> s = v ? v->as_string() : std::nullopt;
  return s ? *s : fallback;
}
double optional_double(const TomlTable* t, std::string_view key, double fallback) {
  const TomlValue* v = lookup(t, key);
  const std::optional<double> d = v ? v->as_double() : std::nullopt;
  return d ? *d : fallback;
}
std::int64_t optional_int(const TomlTable* t, std::string_view key, std::int64_t fallback) {
  const TomlValue* v = lookup(t, key);
  const std::optional<std::int64_t> i = v ? v->as_int() : std::nullopt;
  return i ? *i : fallback;
}
bool optional_bool(const TomlTable* t, std::string_view key, bool fallback) {
  const TomlValue* v = lookup(t, key);
  const std::optional<bool> b = v ? v->as_bool() : std::nullopt;
  return b ? *b : fallback;
}

// One field component: one number for every spin, or one per spin.
std::vector<double> component(const TomlTable* t, std::string_view key, std::size_t nspins,
                              std::string_view origin) {
  std::vector<double> out(nspins, 0.0);
  const TomlValue* v = lookup(t, key);
  if (!v) return out;
  if (const std::optional<double> scalar = v->as_double()) {
    out.assign(nspins, *scalar);
    return out;
  }
  if (v->kind != TomlValue::Kind::Array) throw ConfigError(origin, key, "expected a number or an array of numbers");
  if (v->array.size() != nspins)
    throw ConfigError(origin, key,
                      "has " + std::to_string(v->array.size()) + " entries but there are " +
                          std::to_string(nspins) + " spins");
  for (std::size_t i = 0; i < nspins; ++i) {
    const std::optional<double> d = v->array[i].as_double();
    if (!d) throw ConfigError(origin, key, "contains a non-numeric entry");
    out[i] = *d;
  }
  return out;
}

std::vector<Step> steps_from(const TomlArray& list, std::size_t nspins, const std::vector<Field>& base,
                             std::string_view origin, std::size_t depth);

Step step_from(const TomlTable& t, std::size_t nspins, const std::vector<Field>& base, std::string_view origin,
               std::size_t depth) {
  const std::string kind = require_string(&t, "kind", origin);
  if (kind == "measure") return Measure{};
  if (kind == "evolve") {
    Evolve e;
    e.tau = require_double(&t, "tau", origin);
    const std::int64_t steps = optional_int(&t, "steps", 1);
    if (steps < 0) throw ConfigError(origin, "steps", "must not be negative");
    e.nsteps = static_cast<std::size_t>(steps);
    const std::string mode = optional_string(&t, "mode", "real");
    if (mode == "real")
      e.mode = TimeMode::Real;
    else if (mode == "imaginary")
      e.mode = TimeMode::Imaginary;
    else
      throw ConfigError(origin, "mode", "expected \"real\" or \"imaginary\"");
    e.autonormalize = optional_bool(&t, "autonormalize", false);
    if (t.contains("hx") || t.contains("hy") || t.contains("hz")) {
      // components not named keep the system's baseline
      const std::vector<double> hx = component(&t, "hx", nspins, origin);
      const std::vector<double> hy = component(&t, "hy", nspins, origin);
      const std::vector<double> hz = component(&t, "hz", nspins, origin);
      e.fields.resize(nspins);
      for (std::size_t s = 0; s < nspins; ++s)
        e.fields[s] = Field{t.contains("hx") ? hx[s] : base[s].x, t.contains("hy") ? hy[s] : base[s].y,
                            t.contains("hz") ? hz[s] : base[s].z};
    }
    return e;
  }
  if (kind == "pulse") {
    const std::string spec = require_string(&t, "rotation", origin);
    try {
      return Pulse{makeRotation(spec, nspins)};
    } catch (const std::invalid_argument& e) {
      throw ConfigError(origin, "rotation", e.what());
    }
  }
  if (kind == "repeat") {
    if (depth > 32) throw ConfigError(origin, "repeat", "nested too deeply");
    Repeat r;
    const std::int64_t count = require_int(&t, "count", origin);
    if (count < 0) throw ConfigError(origin, "count", "must not be negative");
    r.count = static_cast<std::size_t>(count);
    const TomlArray* body = array_at(&t, "body");
    if (!body) throw ConfigError(origin, "body", "a repeat needs an array of steps");
    r.body = steps_from(*body, nspins, base, origin, depth + 1);
    return r;
  }
  throw ConfigError(origin, "kind",
                    "unknown step kind '" + kind + "'; expected evolve, pulse, measure or repeat");
}

std::vector<Step> steps_from(const TomlArray& list, std::size_t nspins, const std::vector<Field>& base,
                             std::string_view origin, std::size_t depth) {
  std::vector<Step> out;
  out.reserve(list.size());
  for (const TomlValue& v : list) {
    if (v.kind != TomlValue::Kind::Table) throw ConfigError(origin, "step", "every step must be a table");
    out.push_back(step_from(*v.table, nspins, base, origin, depth));
  }
  return out;
}

}  // namespace

RunConfig parseConfig(std::string_view text, std::string_view origin) {
  const TomlValue root = Reader(text, origin).document();
  const TomlTable& doc = *root.table;

  RunConfig cfg;
  const TomlTable* system = section(doc, "system");
  const std::int64_t spins = require_int(system, "spins", origin);
  if (spins <= 0) throw ConfigError(origin, "system.spins", "must be positive");
  cfg.nspins = static_cast<std::size_t>(spins);
  try {
    cfg.initialState = parseInitialStateKind(optional_string(system, "initial_state", "random"));
  } catch (const std::invalid_argument& e) {
    throw ConfigError(origin, "system.initial_state", e.what());
  }
  cfg.basisIndex = static_cast<std::size_t>(optional_int(system, "basis_index", 0));

  if (cfg.nspins <= kMaxSpins) {
    const TomlTable* field = section(doc, "field");
    const std::vector<double> hx = component(field, "hx", cfg.nspins, origin);
    const std::vector<double> hy = component(field, "hy", cfg.nspins, origin);
    const std::vector<double> hz = component(field, "hz", cfg.nspins, origin);
    cfg.baseFields.resize(cfg.nspins);
    for (std::size_t s = 0; s < cfg.nspins; ++s) cfg.baseFields[s] = Field{hx[s], hy[s], hz[s]};
  }

  const TomlTable* ensemble = section(doc, "ensemble");
  cfg.realizations = static_cast<std::size_t>(optional_int(ensemble, "realizations", 1));
  cfg.seed = static_cast<std::uint64_t>(optional_int(ensemble, "seed", 0));
  cfg.generateGeometry = optional_bool(ensemble, "dipolar_geometry", true);
  cfg.dipolarScale = optional_double(ensemble, "dipolar_scale", 1.0);
  cfg.fieldDisorder = optional_double(ensemble, "field_disorder", 0.0);
  // this build's geometry options (RunConfig::minSeparation, secularAngle);
  // absent, the lattice is the reference's
  cfg.minSeparation = optional_double(ensemble, "min_separation", cfg.minSeparation);
  cfg.secularAngle = optional_bool(ensemble, "secular_angle", cfg.secularAngle);

  const TomlTable* measurement = section(doc, "measurement");
  try {
    cfg.referenceAxis = parseAxis(optional_string(measurement, "reference_axis", "z"));
  } catch (const std::invalid_argument& e) {
    throw ConfigError(origin, "measurement.reference_axis", e.what());
  }
  if (const TomlArray* list = array_at(measurement, "observables")) {
    for (const TomlValue& v : *list) {
      const std::optional<std::string> spec = v.as_string();
      if (!spec) throw ConfigError(origin, "measurement.observables", "entries must be strings");
      cfg.observables.push_back(*spec);
    }
  }

  // [[coupling]] tables: spins = [i, j] (1-based, as in rotation and
  // observable specs), jx, jy, jz (0 when absent)
  if (const TomlArray* list = array_at(&doc, "coupling")) {
    for (const TomlValue& v : *list) {
      if (v.kind != TomlValue::Kind::Table) throw ConfigError(origin, "coupling", "every coupling must be a table");
      const TomlTable* t = v.table.get();
      const TomlArray* spins = array_at(t, "spins");
      std::optional<std::int64_t> a, b;
      if (spins && spins->size() == 2) {
        a = (*spins)[0].as_int();
        b = (*spins)[1].as_int();
      }
      if (!a || !b) throw ConfigError(origin, "coupling.spins", "expected two spin numbers, 1-based");
      const auto n = static_cast<std::int64_t>(cfg.nspins);
      if (*a < 1 || *b < 1 || *a > n || *b > n || *a == *b)
        throw ConfigError(origin, "coupling.spins",
                          "needs two different spins between 1 and " + std::to_string(cfg.nspins));
      cfg.couplings.push_back(Coupling{SpinPair(static_cast<SpinIndex>(*a - 1), static_cast<SpinIndex>(*b - 1)),
                                       optional_double(t, "jx", 0.0), optional_double(t, "jy", 0.0),
                                       optional_double(t, "jz", 0.0)});
    }
  }

  cfg.epsilon = optional_double(section(doc, "numerics"), "epsilon", 1.0e-7);

  const TomlArray* steps = array_at(&doc, "step");
  if (!steps) throw ConfigError(origin, "step", "missing; a run needs a [[step]] list");
  if (cfg.nspins <= kMaxSpins) cfg.sequence = Sequence(steps_from(*steps, cfg.nspins, cfg.baseFields, origin, 0));

  cfg.validate(origin);
  return cfg;
}

RunConfig loadConfig(const std::filesystem::path& path) {
  std::ifstream in(path);
  if (!in) throw ConfigError(path.string(), "<file>", "cannot be opened for reading");
  const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
  return parseConfig(text, path.string());
}

}  // namespace spinsim
