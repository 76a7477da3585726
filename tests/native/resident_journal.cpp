// Line-driven harness for csrc/resident_journal.hpp (built with g++ by
// tests/test_resident_journal.py, which holds the brute-force model).
//   begin | close | log
//   write <k> (<slot> <lo> <hi> <width>){k}   plan, then record the saves
//   transpose <rows> <cols> <start> <n>
//   handle <tag> <slot> | slot <handle> <tag> <slots>
#include <iostream>
#include <sstream>
#include <string>

#include "resident_journal.hpp"

using namespace dvgpu::resident;

int main(int argc, char** argv) {
    Journal j(argc > 1 ? std::stoull(argv[1]) : 0);
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string cmd;
        in >> cmd;
        std::ostringstream out;
        try {
            if (cmd == "begin") {
                j.begin();
                out << "ok";
            } else if (cmd == "close") {
                j.close();
                out << "ok";
            } else if (cmd == "log") {
                out << "ok " << j.used() << " " << j.log().size();
                for (const SavedRun& s : j.log())
                    out << " " << s.slot << " " << s.lo << " " << s.count << " " << s.width << " " << s.arena;
            } else if (cmd == "write") {
                size_t k = 0;
                in >> k;
                std::vector<WriteRange> w(k);
                for (auto& r : w) in >> r.slot >> r.lo >> r.hi >> r.width;
                const auto saves = j.plan(w);
                j.record(saves);
                out << "ok " << saves.size();
                for (const SavedRun& s : saves)
                    out << " " << s.slot << " " << s.lo << " " << s.count << " " << s.width << " " << s.arena;
            } else if (cmd == "transpose") {
                uint64_t rows, cols, start, n;
                in >> rows >> cols >> start >> n;
                std::vector<Run> runs;
                transpose_destinations(rows, cols, start, n, runs);
                out << "ok " << runs.size();
                for (const Run& r : runs) out << " " << r.lo << " " << r.hi;
            } else if (cmd == "handle") {
                uint32_t tag, slot;
                in >> tag >> slot;
                out << "ok " << handle_of(tag, slot);
            } else if (cmd == "slot") {
                uint64_t h;
                uint32_t tag;
                size_t slots;
                in >> h >> tag >> slots;
                out << "ok " << slot_of(h, tag, slots);
            } else {
                out << "err unknown command";
            }
        } catch (const dvgpu::InvalidArgument& e) {
            out.str("");
            out << "err " << e.what();
        }
        std::cout << out.str() << std::endl;
    }
    return 0;
}
