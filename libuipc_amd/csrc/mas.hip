// MAS preconditioner (placeholder until the engine lands).
#include "linsys.hpp"

namespace mi
{
std::unique_ptr<LocalPreconditioner> make_mas(const mi_precond_desc&)
{
    throw Error(MI_ERR_UNSUPPORTED, "MAS preconditioner not built yet");
}
}  // namespace mi
