// Minimal XML DOM reader for model.xml / cross_sections.xml.
//
// The reference parses its inputs with pugixml (src/xml_interface.cpp), which
// is a vendored submodule and absent here.  This reader covers the XML the
// model files use: prolog, comments, elements, quoted attributes, text and the
// five predefined entities.  Values are looked up the way the reference's
// get_node_value does (src/xml_interface.cpp): an attribute of that name if
// present, otherwise the text of a child element.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace omc {

struct XmlNode {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::string text;
  std::vector<std::unique_ptr<XmlNode>> children;

  const XmlNode* child(const std::string& n) const
  {
    for (const auto& c : children)
      if (c->name == n)
        return c.get();
    return nullptr;
  }
  std::vector<const XmlNode*> children_named(const std::string& n) const
  {
    std::vector<const XmlNode*> out;
    for (const auto& c : children)
      if (c->name == n)
        out.push_back(c.get());
    return out;
  }
  const std::string* attr(const std::string& n) const
  {
    for (const auto& a : attrs)
      if (a.first == n)
        return &a.second;
    return nullptr;
  }
  // check_for_node (src/xml_interface.cpp)
  bool has(const std::string& n) const { return attr(n) != nullptr || child(n) != nullptr; }
  // get_node_value (src/xml_interface.cpp), whitespace trimmed
  std::string value(const std::string& n) const;
};

std::unique_ptr<XmlNode> parse_xml_file(const std::string& path);
std::unique_ptr<XmlNode> parse_xml_string(const std::string& text);

std::vector<std::string> split_ws(const std::string& s);
std::vector<double> to_doubles(const std::string& s);
std::vector<long long> to_ints(const std::string& s);
std::string trim(const std::string& s);

} // namespace omc
