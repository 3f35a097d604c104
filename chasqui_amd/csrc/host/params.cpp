#include "params.hpp"

#include <algorithm>
#include <cctype>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace chq {

namespace {
std::string strip(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
  while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
  return s.substr(a, b - a);
}
std::string full_key(const std::string &b, const std::string &k) { return b + "." + k; }
}  // namespace

void Params::read_text(const std::string &text) {
  std::istringstream in(text);
  std::string line, block;
  while (std::getline(in, line)) {
    const auto hash = line.find('#');
    if (hash != std::string::npos) line.erase(hash);
    line = strip(line);
    if (line.empty()) continue;
    if (line.front() == '[' && line.back() == ']') {
      block = strip(line.substr(1, line.size() - 2));
      continue;
    }
    const auto eq = line.find('=');
    if (eq == std::string::npos) continue;
    table_[full_key(block, strip(line.substr(0, eq)))] = strip(line.substr(eq + 1));
  }
}

void Params::read_file(const std::string &path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open input file: " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  read_text(ss.str());
}

void Params::override_kv(const std::string &arg) {
  const auto eq = arg.find('=');
  if (eq == std::string::npos) throw std::runtime_error("bad override: " + arg);
  table_[strip(arg.substr(0, eq))] = strip(arg.substr(eq + 1));
}

bool Params::has(const std::string &b, const std::string &k) const { return table_.count(full_key(b, k)) > 0; }

std::string Params::get_string(const std::string &b, const std::string &k) const {
  auto it = table_.find(full_key(b, k));
  if (it == table_.end()) throw std::runtime_error("missing parameter " + full_key(b, k));
  return it->second;
}

std::string Params::get_string(const std::string &b, const std::string &k, const std::string &d) {
  auto r = table_.emplace(full_key(b, k), d);
  return r.first->second;
}

double Params::get_real(const std::string &b, const std::string &k) const { return std::stod(get_string(b, k)); }

double Params::get_real(const std::string &b, const std::string &k, double d) {
  if (!has(b, k)) {
    std::ostringstream o;
    o.precision(17);   // round-trips every double exactly
    o << d;
    table_[full_key(b, k)] = o.str();
  }
  return std::stod(table_[full_key(b, k)]);
}

int Params::get_int(const std::string &b, const std::string &k) const { return std::stoi(get_string(b, k)); }

int Params::get_int(const std::string &b, const std::string &k, int d) {
  if (!has(b, k)) table_[full_key(b, k)] = std::to_string(d);
  return std::stoi(table_[full_key(b, k)]);
}

bool Params::get_bool(const std::string &b, const std::string &k, bool d) {
  if (!has(b, k)) table_[full_key(b, k)] = d ? "true" : "false";
  std::string v = table_[full_key(b, k)];
  std::transform(v.begin(), v.end(), v.begin(), [](unsigned char c) { return char(std::tolower(c)); });
  return v == "true" || v == "1" || v == "yes";
}

void Params::print(std::ostream &os) const {
  std::string last;
  for (const auto &kv : table_) {
    const auto dot = kv.first.find('.');
    const std::string b = kv.first.substr(0, dot), k = kv.first.substr(dot + 1);
    if (b != last) {
      os << "[" << b << "]\n";
      last = b;
    }
    os << "  " << k << " = " << kv.second << "\n";
  }
}

}  // namespace chq
