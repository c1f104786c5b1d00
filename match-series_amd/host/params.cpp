#include "params.h"

#include <sys/stat.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <set>
#include <sstream>

namespace mshost {

void Exception::dump(std::ostream &os) const {
  os << "\n--------------------------------------------------------------------------------\n"
     << "Exception in " << file_ << ", line " << line_ << ":\n"
     << what() << "\n"
     << "--------------------------------------------------------------------------------\n";
}

std::string strprintf(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  char buf[4096];
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return buf;
}

std::string expandTilde(const std::string &s) {
  if (!s.empty() && s[0] == '~') {
    const char *home = getenv("HOME");
    return std::string(home ? home : "") + s.substr(1);
  }
  return s;
}

void makeDirectory(const std::string &dir) {
  std::string cur;
  for (size_t i = 0; i < dir.size(); ++i) {
    cur += dir[i];
    if (dir[i] == '/' || i + 1 == dir.size()) mkdir(cur.c_str(), 0755);
  }
}

bool fileExists(const std::string &f) {
  struct stat st;
  return stat(f.c_str(), &st) == 0;
}

namespace {
bool ws(int c) { return c == ' ' || c == '\t' || c == '\r'; }

void skipWs(std::istream &in) {
  while (!in.eof() && ws(in.peek())) in.get();
}

// "{ a b { c d } }" -> flat list (nesting only groups values)
void readList(std::istream &in, std::vector<std::string> &out) {
  int depth = 0;
  std::string tok;
  while (!in.eof()) {
    const int c = in.get();
    if (c == EOF) break;
    if (c == '{') {
      ++depth;
      continue;
    }
    if (c == '}' || ws(c) || c == '\n') {
      if (!tok.empty()) {
        out.push_back(tok);
        tok.clear();
      }
      if (c == '}' && --depth == 0) return;
      continue;
    }
    tok += (char)c;
  }
  MSH_THROW("Unexpected end of file while parsing a list");
}
}  // namespace

ParameterParser::ParameterParser(const std::string &fileName) : file_(fileName) {
  std::ifstream in(fileName);
  if (!in) MSH_THROW(strprintf("Could not open parameter file \"%s\"", fileName.c_str()));
  while (true) {
    skipWs(in);
    while (!in.eof() && (in.peek() == '#' || in.peek() == '\n' || ws(in.peek()))) {
      if (ws(in.peek())) skipWs(in);
      else {
        std::string dummy;
        std::getline(in, dummy);
      }
    }
    if (in.eof() || in.peek() == EOF) break;
    Var v;
    in >> v.name;
    if (hasVariable(v.name)) MSH_THROW(strprintf("Variable \"%s\" already defined", v.name.c_str()));
    skipWs(in);
    if (in.eof() || in.peek() == '\n' || in.peek() == EOF)
      MSH_THROW(strprintf("Unexpected end of line while parsing variable %s", v.name.c_str()));
    if (in.peek() == '{') {
      v.list = true;
      readList(in, v.values);
    } else if (in.peek() == '"') {
      in.get();
      std::string full;
      while (!in.eof() && in.peek() != '"' && in.peek() != '\n' && in.peek() != EOF) full += (char)in.get();
      if (in.peek() != '"')
        MSH_THROW(strprintf("Unexpected end of line while parsing quotation mark enclosed variable %s", v.name.c_str()));
      in.get();
      v.values.push_back(full);
    } else {
      std::string val;
      in >> val;
      v.values.push_back(val);
    }
    vars_.push_back(v);
    skipWs(in);
    if (!in.eof() && in.peek() != '\n' && in.peek() != EOF)
      MSH_THROW(strprintf("Error while parsing variable \"%s\": Unexpected data after the variable value found",
                          v.name.c_str()));
  }
}

bool ParameterParser::hasVariable(const std::string &name) const {
  for (const auto &v : vars_)
    if (v.name == name) return true;
  return false;
}

const ParameterParser::Var &ParameterParser::find(const std::string &name) const {
  for (const auto &v : vars_)
    if (v.name == name) return v;
  MSH_THROW(strprintf("No match found for %s.\n", name.c_str()));
}

ParameterParser::Var *ParameterParser::findMutable(const std::string &name) {
  for (auto &v : vars_)
    if (v.name == name) return &v;
  return nullptr;
}

std::string ParameterParser::getString(const std::string &name, int index) const {
  const Var &v = find(name);
  if (index < 0 || index >= (int)v.values.size())
    MSH_THROW(strprintf("Index %d out of range for variable %s", index, name.c_str()));
  return v.values[index];
}

int ParameterParser::getInt(const std::string &name, int index) const {
  return atoi(getString(name, index).c_str());
}

double ParameterParser::getDouble(const std::string &name, int index) const {
  return atof(getString(name, index).c_str());
}

std::string ParameterParser::getStringExpandTilde(const std::string &name) const {
  return expandTilde(getString(name));
}

int ParameterParser::getIntOrDefault(const std::string &name, int def) const {
  return hasVariable(name) ? getInt(name) : def;
}

double ParameterParser::getDoubleOrDefault(const std::string &name, double def) const {
  return hasVariable(name) ? getDouble(name) : def;
}

bool ParameterParser::checkAndGetBool(const std::string &name) const {
  return hasVariable(name) && getInt(name) == 1;
}

int ParameterParser::getDimSize(const std::string &name) const { return (int)find(name).values.size(); }

void ParameterParser::changeVariableValue(const std::string &name, const std::string &value) {
  Var *v = findMutable(name);
  if (!v) MSH_THROW(strprintf("Cannot change the value of the unknown variable %s", name.c_str()));
  v->values.assign(1, value);
  v->list = false;
}

void ParameterParser::changeVariableValue(const std::string &name, int value) {
  changeVariableValue(name, strprintf("%d", value));
}

void ParameterParser::addVariable(const std::string &name, const std::string &value) {
  if (hasVariable(name)) MSH_THROW(strprintf("Variable \"%s\" already defined", name.c_str()));
  Var v;
  v.name = name;
  v.values.push_back(value);
  vars_.push_back(v);
}

void ParameterParser::addVariable(const std::string &name, int value) { addVariable(name, strprintf("%d", value)); }

void ParameterParser::dump(std::ostream &os) const {
  for (const auto &v : vars_) {
    os << v.name << " ";
    if (v.list) {
      os << "{ ";
      for (const auto &x : v.values) os << x << " ";
      os << "}";
    } else {
      const std::string &x = v.values[0];
      if (x.find(' ') != std::string::npos || x.empty()) os << '"' << x << '"';
      else os << x;
    }
    os << "\n";
  }
}

void ParameterParser::dumpToFile(const std::string &fileName, const std::string &directory) const {
  std::ofstream out(directory + fileName);
  if (!file_.empty()) out << "# Initialized from file " << file_ << "\n";
  dump(out);
}

std::vector<std::string> createTemplateFileNameList(const ParameterParser &p) {
  std::set<int> skip;
  if (p.hasVariable("templateSkipNums"))
    for (int i = 0; i < p.getDimSize("templateSkipNums"); ++i) skip.insert(p.getInt("templateSkipNums", i));
  std::vector<std::string> names;
  if (p.hasVariable("numTemplates")) {
    const int n = p.getInt("numTemplates");
    const std::string pattern = p.getStringExpandTilde("templateNamePattern");
    for (int i = 0; i < n; ++i) {
      const int index = p.getInt("templateNumStep") * i + p.getInt("templateNumOffset");
      if (skip.count(index)) continue;
      names.push_back(strprintf(pattern.c_str(), index));
    }
  } else {
    const std::string dir = p.hasVariable("templatesDirectory") ? p.getString("templatesDirectory") : "";
    for (int i = 0; i < p.getDimSize("templates"); ++i) names.push_back(dir + p.getString("templates", i));
  }
  return names;
}

}  // namespace mshost
