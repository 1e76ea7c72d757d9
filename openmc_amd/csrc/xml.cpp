#include "xml.h"

#include <cstdlib>
#include <fstream>
#include <sstream>

namespace omc {

std::string trim(const std::string& s)
{
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos)
    return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

std::vector<std::string> split_ws(const std::string& s)
{
  std::vector<std::string> out;
  std::istringstream in(s);
  std::string w;
  while (in >> w)
    out.push_back(w);
  return out;
}

std::vector<double> to_doubles(const std::string& s)
{
  std::vector<double> out;
  for (const auto& w : split_ws(s)) {
    char* end = nullptr;
    double v = std::strtod(w.c_str(), &end);
    if (end == w.c_str() || *end != '\0')
      throw std::runtime_error("not a number: '" + w + "'");
    out.push_back(v);
  }
  return out;
}

std::vector<long long> to_ints(const std::string& s)
{
  std::vector<long long> out;
  for (const auto& w : split_ws(s)) {
    char* end = nullptr;
    long long v = std::strtoll(w.c_str(), &end, 10);
    if (end == w.c_str() || *end != '\0')
      throw std::runtime_error("not an integer: '" + w + "'");
    out.push_back(v);
  }
  return out;
}

std::string XmlNode::value(const std::string& n) const
{
  if (const std::string* a = attr(n))
    return trim(*a);
  if (const XmlNode* c = child(n))
    return trim(c->text);
  throw std::runtime_error("missing XML value '" + n + "' in <" + name + ">");
}

namespace {

struct Parser {
  const std::string& s;
  size_t i = 0;

  explicit Parser(const std::string& src) : s(src) {}

  [[noreturn]] void fail(const std::string& msg)
  {
    throw std::runtime_error("XML parse error at offset " + std::to_string(i) + ": " + msg);
  }

  bool starts(const char* t) const { return s.compare(i, std::char_traits<char>::length(t), t) == 0; }

  void skip_ws()
  {
    while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\r' || s[i] == '\n'))
      ++i;
  }

  void skip_misc()
  {
    while (true) {
      skip_ws();
      if (starts("<?")) {
        size_t e = s.find("?>", i);
        if (e == std::string::npos)
          fail("unterminated processing instruction");
        i = e + 2;
      } else if (starts("<!--")) {
        size_t e = s.find("-->", i);
        if (e == std::string::npos)
          fail("unterminated comment");
        i = e + 3;
      } else if (starts("<!DOCTYPE")) {
        size_t e = s.find('>', i);
        if (e == std::string::npos)
          fail("unterminated doctype");
        i = e + 1;
      } else {
        break;
      }
    }
  }

  static std::string decode(const std::string& t)
  {
    std::string out;
    out.reserve(t.size());
    for (size_t k = 0; k < t.size(); ++k) {
      if (t[k] == '&') {
        size_t e = t.find(';', k);
        if (e == std::string::npos)
          throw std::runtime_error("XML parse error: '&' without a closing ';' in \"" + t + "\"");
        std::string ent = t.substr(k + 1, e - k - 1);
        if (ent == "lt")
          out += '<';
        else if (ent == "gt")
          out += '>';
        else if (ent == "amp")
          out += '&';
        else if (ent == "quot")
          out += '"';
        else if (ent == "apos")
          out += '\'';
        else
          out += "&" + ent + ";";
        k = e;
      } else {
        out += t[k];
      }
    }
    return out;
  }

  std::string read_name()
  {
    size_t a = i;
    while (i < s.size() && !(s[i] == ' ' || s[i] == '\t' || s[i] == '\r' || s[i] == '\n' ||
                             s[i] == '/' || s[i] == '>' || s[i] == '='))
      ++i;
    if (a == i)
      fail("expected a name");
    return s.substr(a, i - a);
  }

  std::unique_ptr<XmlNode> element()
  {
    if (s[i] != '<')
      fail("expected '<'");
    ++i;
    auto node = std::make_unique<XmlNode>();
    node->name = read_name();
    while (true) {
      skip_ws();
      if (i >= s.size())
        fail("unterminated start tag");
      if (starts("/>")) {
        i += 2;
        return node;
      }
      if (s[i] == '>') {
        ++i;
        break;
      }
      std::string an = read_name();
      skip_ws();
      if (s[i] != '=')
        fail("expected '=' after attribute " + an);
      ++i;
      skip_ws();
      char q = s[i];
      if (q != '"' && q != '\'')
        fail("expected quoted attribute value");
      size_t e = s.find(q, i + 1);
      if (e == std::string::npos)
        fail("unterminated attribute value");
      node->attrs.emplace_back(an, decode(s.substr(i + 1, e - i - 1)));
      i = e + 1;
    }
    // content
    while (true) {
      if (i >= s.size())
        fail("unterminated element <" + node->name + ">");
      if (starts("</")) {
        i += 2;
        std::string en = read_name();
        if (en != node->name)
          fail("mismatched end tag </" + en + "> for <" + node->name + ">");
        skip_ws();
        if (s[i] != '>')
          fail("expected '>'");
        ++i;
        return node;
      } else if (starts("<!--")) {
        size_t e = s.find("-->", i);
        if (e == std::string::npos)
          fail("unterminated comment");
        i = e + 3;
      } else if (starts("<![CDATA[")) {
        size_t e = s.find("]]>", i);
        node->text += s.substr(i + 9, e - i - 9);
        i = e + 3;
      } else if (s[i] == '<') {
        node->children.push_back(element());
      } else {
        size_t e = s.find('<', i);
        if (e == std::string::npos)
          fail("unterminated text");
        node->text += decode(s.substr(i, e - i));
        i = e;
      }
    }
  }
};

} // namespace

std::unique_ptr<XmlNode> parse_xml_string(const std::string& text)
{
  Parser p(text);
  p.skip_misc();
  if (p.i >= text.size())
    throw std::runtime_error("empty XML document");
  return p.element();
}

std::unique_ptr<XmlNode> parse_xml_file(const std::string& path)
{
  std::ifstream f(path);
  if (!f)
    throw std::runtime_error("cannot open " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return parse_xml_string(ss.str());
}

} // namespace omc
