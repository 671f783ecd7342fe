// serialize.hpp -- binary serialization of the host API objects.
//
// Mirrors FIDESlib's fideslib::Serial (api/Serialize.hpp:10-27) and the key
// (de)serialization of CryptoContext (api/CryptoContext.hpp:90-95).  The
// reference writes OpenFHE's cereal formats plus a ".dev" sidecar
// (api/Serialize.cpp:14-208); OpenFHE is not available here, so the format
// is this library's own:
//
//   "FIDESAMD" | u32 version (1) | u32 object tag | payload
//
// little-endian, u64 integers and IEEE doubles, strings as u64 length +
// bytes.  Payloads: parameters (ring, moduli, roots, digits, technique,
// scaling tables), ciphertext (limbs, metadata, [2][limbs][N] residues in
// evaluation form), key-switching key ([2][dnum][L+K][N]), context
// (parameters + every loaded key with its automorphism index).  Only BINARY
// is supported; JSON returns false.
#ifndef FIDESLIB_AMD_SERIALIZE_HPP
#define FIDESLIB_AMD_SERIALIZE_HPP

#include <iosfwd>
#include <string>

#include "fideslib/fideslib.hpp"

namespace fideslib {

enum SerType { BINARY = 0, JSON = 1 };

namespace Serial {

bool Serialize(std::ostream& os, const RawParams& obj, SerType t = BINARY);
bool Serialize(std::ostream& os, const RawCipherText& obj, SerType t = BINARY);
bool Serialize(std::ostream& os, const RawKeySwitchKey& obj, SerType t = BINARY);
bool Serialize(std::ostream& os, const CryptoContext& cc, SerType t = BINARY);   // parameters + loaded keys
bool Deserialize(std::istream& is, RawParams& obj, SerType t = BINARY);
bool Deserialize(std::istream& is, RawCipherText& obj, SerType t = BINARY);
bool Deserialize(std::istream& is, RawKeySwitchKey& obj, SerType t = BINARY);
bool Deserialize(std::istream& is, CryptoContext& cc, SerType t = BINARY);       // creates the context, uploads keys

template <class T>
bool SerializeToFile(const std::string& filename, const T& obj, SerType t = BINARY);
template <class T>
bool DeserializeFromFile(const std::string& filename, T& obj, SerType t = BINARY);

}  // namespace Serial
}  // namespace fideslib

#endif
