// parse_string: split a line on a (multi-character) delimiter, keeping empty
// fields -- the helper GraphHET's LDBC loader and checkers call
// (include/CPU_adj_list/parse_string.hpp).  "a = b" split on " = " gives
// {"a", "b"}; an empty delimiter returns the whole line as one field.
#pragma once

#include <string>
#include <vector>

inline std::vector<std::string> parse_string(const std::string& text, const std::string& delimiter) {
    std::vector<std::string> fields;
    if (delimiter.empty()) {
        fields.push_back(text);
        return fields;
    }
    std::size_t from = 0;
    for (;;) {
        const std::size_t at = text.find(delimiter, from);
        if (at == std::string::npos) break;
        fields.emplace_back(text, from, at - from);
        from = at + delimiter.size();
    }
    fields.emplace_back(text, from, std::string::npos);
    return fields;
}
