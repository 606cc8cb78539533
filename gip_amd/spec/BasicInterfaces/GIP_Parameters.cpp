#include "GIP_Parameters.h"

std::ostream &operator<<(std::ostream &q, const GIP_Parameters &p) {
    q << "   START PARAMETERS " << std::endl << std::endl << "    -Booleans:" << std::endl;
    for (const auto &kv : p.bools) q << "     " << kv.first << ": " << kv.second << std::endl;
    q << std::endl << "    -Integers:" << std::endl;
    for (const auto &kv : p.integers) q << "     " << kv.first << ": " << kv.second << std::endl;
    q << std::endl << "    -Doubles:" << std::endl;
    for (const auto &kv : p.doubles) q << "     " << kv.first << ": " << kv.second << std::endl;
    q << std::endl << "    -Strings:" << std::endl;
    for (const auto &kv : p.strings) q << "     " << kv.first << ": " << kv.second << std::endl;
    q << std::endl << "   END PARAMETERS" << std::endl;
    return q;
}
