// dacti/driver/io/htg_writer.hpp -- io::htg_writer for callers of the
// reference's hyper-tree-grid writer (driver/io/htg_writer.hpp:15-196).
//
// The reference writes <name>_<frame>.htg through VTK's
// vtkXMLHyperTreeGridWriter; VTK is not part of this build, so htg_writer
// keeps the constructor and write_data() interface and writes the same
// frame as io::vtu_writer does (<name>_<frame>.vtu: the active leaves with
// their primitives, parameters and metadata).
#ifndef DACTI_HTG_WRITER_HPP
#define DACTI_HTG_WRITER_HPP

#include "vtu_writer.hpp"

namespace dacti::io {

class htg_writer : public vtu_writer {
public:
    using vtu_writer::vtu_writer;
};

}  // namespace dacti::io

#endif
