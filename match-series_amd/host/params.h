// Parameter files of matchSeries (the reference's ".par" surface) and the errors of the host layer.
#pragma once

#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

namespace mshost {

// Error carrying the source location, reported like the reference's aol::Exception::dump
// (aol/qmException.h:93-96).
class Exception : public std::runtime_error {
 public:
  Exception(const std::string &msg, const char *file, int line)
      : std::runtime_error(msg), file_(file), line_(line) {}
  void dump(std::ostream &os) const;

 private:
  std::string file_;
  int line_;
};
#define MSH_THROW(msg) throw ::mshost::Exception((msg), __FILE__, __LINE__)

std::string strprintf(const char *fmt, ...);

// A parameter file: one "name value" per line, '#' comments, quoted strings with spaces, and
// "{ ... }" lists (possibly nested), every name at most once (aol/parameterParser.cpp:154-219).
class ParameterParser {
 public:
  ParameterParser() = default;
  explicit ParameterParser(const std::string &fileName);

  bool hasVariable(const std::string &name) const;
  int getInt(const std::string &name, int index = 0) const;
  double getDouble(const std::string &name, int index = 0) const;
  std::string getString(const std::string &name, int index = 0) const;
  std::string getStringExpandTilde(const std::string &name) const;
  int getIntOrDefault(const std::string &name, int def) const;
  double getDoubleOrDefault(const std::string &name, double def) const;
  // true iff the variable exists and equals 1 (parameterParser.cpp:436-438)
  bool checkAndGetBool(const std::string &name) const;
  int getDimSize(const std::string &name) const;

  void changeVariableValue(const std::string &name, const std::string &value);
  void changeVariableValue(const std::string &name, int value);
  void addVariable(const std::string &name, const std::string &value);
  void addVariable(const std::string &name, int value);

  void dump(std::ostream &os) const;
  void dumpToFile(const std::string &fileName, const std::string &directory) const;
  const std::string &fileName() const { return file_; }

 private:
  struct Var {
    std::string name;
    std::vector<std::string> values;
    bool list = false;
  };
  const Var &find(const std::string &name) const;
  Var *findMutable(const std::string &name);
  std::vector<Var> vars_;
  std::string file_;
};

// templateNamePattern / templateNumOffset / templateNumStep / numTemplates / templateSkipNums, or
// the "templates" list (aol/parameterParser.cpp:508-541).
std::vector<std::string> createTemplateFileNameList(const ParameterParser &p);

std::string expandTilde(const std::string &s);
void makeDirectory(const std::string &dir);
bool fileExists(const std::string &f);

}  // namespace mshost
