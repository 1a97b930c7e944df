#ifndef IRLBA_HIP_MATRIX_INTERFACE_HPP
#define IRLBA_HIP_MATRIX_INTERFACE_HPP

// The matrix plugin interface of the reference (include/irlba/Matrix/interface.hpp:23-223):
// a Matrix reports its dimensions and creates workspaces that multiply by a
// vector, multiply by the adjoint, and realize a dense copy.  Class names,
// method names and argument roles are kept so that user subclasses and the
// reference's decorators port unchanged.

#include <memory>

#include "../types.hpp"

struct irlba_hip_matrix;

namespace irlba {

// Workspace::multiply (interface.hpp:49): output = A * right; output is
// pre-sized by the caller and fully overwritten.
template <class EigenVector_>
class Workspace {
public:
    virtual ~Workspace() = default;
    virtual void multiply(const EigenVector_& right, EigenVector_& output) = 0;
};

// AdjointWorkspace::multiply (interface.hpp:87): output = A' * right.
template <class EigenVector_>
class AdjointWorkspace {
public:
    virtual ~AdjointWorkspace() = default;
    virtual void multiply(const EigenVector_& right, EigenVector_& output) = 0;
};

// RealizeWorkspace (interface.hpp:98-131).
template <class EigenMatrix_>
class RealizeWorkspace {
public:
    virtual ~RealizeWorkspace() = default;
    virtual const EigenMatrix_& realize(EigenMatrix_& buffer) = 0;
    void realize_copy(EigenMatrix_& buffer) {
        const auto& out = realize(buffer);
        if (&out != &buffer) buffer = out;
    }
};

// Matrix (interface.hpp:141-223).
template <class EigenVector_, class EigenMatrix_>
class Matrix {
public:
    virtual ~Matrix() = default;
    virtual Index rows() const = 0;
    virtual Index cols() const = 0;
    virtual std::unique_ptr<Workspace<EigenVector_> > new_workspace() const = 0;
    virtual std::unique_ptr<AdjointWorkspace<EigenVector_> > new_adjoint_workspace() const = 0;
    virtual std::unique_ptr<RealizeWorkspace<EigenMatrix_> > new_realize_workspace() const = 0;
    std::unique_ptr<Workspace<EigenVector_> > new_known_workspace() const { return new_workspace(); }
    std::unique_ptr<AdjointWorkspace<EigenVector_> > new_known_adjoint_workspace() const {
        return new_adjoint_workspace();
    }
    std::unique_ptr<RealizeWorkspace<EigenMatrix_> > new_known_realize_workspace() const {
        return new_realize_workspace();
    }
};

// Mixin of matrices whose data and decorators live on the GPU: compute() and
// pca() run such a matrix entirely on the device through its handle.
class DeviceResident {
public:
    virtual ~DeviceResident() = default;
    virtual const irlba_hip_matrix* device_handle() const = 0;
};

}  // namespace irlba

#endif
