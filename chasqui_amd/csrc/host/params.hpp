// Input decks: INI blocks `[block]`, lines `key = value`, `#` comments, plus command-line
// `block.key=value` overrides applied after the file (params.hpp:17-44, main.cpp:106-110).
// Typed getters with a default record that default into the table, so printing the table
// shows the effective configuration (params.hpp:48-81, 90).
#pragma once
#include <map>
#include <ostream>
#include <string>

namespace chq {

class Params {
 public:
  void read_text(const std::string &text);
  void read_file(const std::string &path);
  void override_kv(const std::string &arg);   // "block.key=value"
  bool has(const std::string &block, const std::string &key) const;
  std::string get_string(const std::string &block, const std::string &key) const;   // throws if absent
  std::string get_string(const std::string &block, const std::string &key, const std::string &dflt);
  double get_real(const std::string &block, const std::string &key) const;
  double get_real(const std::string &block, const std::string &key, double dflt);
  int get_int(const std::string &block, const std::string &key) const;
  int get_int(const std::string &block, const std::string &key, int dflt);
  bool get_bool(const std::string &block, const std::string &key, bool dflt);
  void print(std::ostream &os) const;

 private:
  std::map<std::string, std::string> table_;
};

}  // namespace chq
