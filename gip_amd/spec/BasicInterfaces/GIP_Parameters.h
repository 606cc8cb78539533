// Typed key/value store (interface of GIP_Parameters.h:12-47 in the
// reference).  Unlike the reference, reading a missing key throws
// std::out_of_range instead of dereferencing end().
#ifndef _gip_parameters_
#define _gip_parameters_

#include <map>
#include <ostream>
#include <string>

class GIP_Parameters {
    friend std::ostream &operator<<(std::ostream &, const GIP_Parameters &);

   public:
    GIP_Parameters() {}

    bool getBool(const std::string &key) const { return bools.at(key); }
    int getInt(const std::string &key) const { return integers.at(key); }
    const double &getDouble(const std::string &key) const { return doubles.at(key); }
    const std::string &getString(const std::string &key) const { return strings.at(key); }
    const char *getChar(const std::string &key) const { return strings.at(key).c_str(); }

    bool hasBool(const std::string &k) const { return bools.count(k) != 0; }
    bool hasInt(const std::string &k) const { return integers.count(k) != 0; }
    bool hasDouble(const std::string &k) const { return doubles.count(k) != 0; }
    bool hasString(const std::string &k) const { return strings.count(k) != 0; }

    void insertBool(const std::string &key, bool value) { bools[key] = value; }
    void insertInt(const std::string &key, int value) { integers[key] = value; }
    void insertDouble(const std::string &key, const double &value) { doubles[key] = value; }
    void insertString(const std::string &key, const std::string &value) { strings[key] = value; }
    void setBool(const std::string &key, bool value) { bools[key] = value; }
    void setInt(const std::string &key, int value) { integers[key] = value; }
    void setDouble(const std::string &key, const double &value) { doubles[key] = value; }
    void setString(const std::string &key, const std::string &value) { strings[key] = value; }

   protected:
    std::map<std::string, bool> bools;
    std::map<std::string, int> integers;
    std::map<std::string, double> doubles;
    std::map<std::string, std::string> strings;
};

#endif
