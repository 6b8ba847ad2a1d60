// Mechanism-tiled forms (FP64 and mixed precision).
#include "forms.hpp"

namespace combustion {

std::string tiled_source(const ProgramView&, bool, unsigned) { throw Unsupported("tiled forms are not built yet"); }
void compile_tiled(const ProgramView&, bool, unsigned) { throw Unsupported("tiled forms are not built yet"); }
std::unique_ptr<Form> assemble_tiled(const ProgramView&, bool, unsigned) { throw Unsupported("tiled forms are not built yet"); }

}  // namespace combustion
