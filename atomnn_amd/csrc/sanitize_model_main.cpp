// The host library's potential-file parser and device-plan builder (model.cpp) on each file given,
// for a build with AddressSanitizer and UndefinedBehaviorSanitizer (make sanitize).  A malformed
// file must end in the parser's own exception (reported, exit status unaffected); memory errors
// and undefined behaviour abort through the sanitizers.
// Usage: sanitize_model <potential.ann>...   Prints one line per file.
#include <cstdio>
#include <exception>

#include "model.hpp"

int main(int argc, char** argv) {
    int parsed = 0, rejected = 0;
    for (int a = 1; a < argc; ++a) {
        try {
            const atomnn::HostModel h = atomnn::parse_ann(argv[a]);
            const atomnn::PlanBuffers pb = atomnn::build_plan(h);
            std::printf("parsed %s: %d species, ldg %d\n", argv[a], h.T, pb.plan.ldg);
            ++parsed;
        } catch (const std::exception& e) {
            std::printf("rejected %s: %s\n", argv[a], e.what());
            ++rejected;
        }
    }
    std::printf("sanitize_model: %d parsed, %d rejected\n", parsed, rejected);
    return 0;
}
