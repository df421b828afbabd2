!> abstract_iterative_solver_type of components/science/source/solver/
!> sci_iterative_solver_mod.f90:38-57 and, for building the shim without LFRic, its
!> conjugate_gradient_type (:87-121, cg_solve :495-592), bicgstab_type (:123-156,
!> bicgstab_solve :645-779) and gmres_type (:161-196, gmres_solve :830-1003) in the same
!> operation order, written against the abstract types only.
module sci_iterative_solver_mod
  use constants_mod,          only : r_def, i_def, l_def, EPS
  use vector_mod,             only : abstract_vector_type
  use sci_linear_operator_mod, only : abstract_linear_operator_type
  use sci_preconditioner_mod, only : abstract_preconditioner_type
  use log_mod,                only : log_event, log_scratch_space, LOG_LEVEL_INFO, &
                                     LOG_LEVEL_ERROR, LOG_LEVEL_DEBUG
  use, intrinsic :: ieee_arithmetic, only : ieee_is_nan
  implicit none

  type, public, abstract :: abstract_iterative_solver_type
    private
    class(abstract_preconditioner_type),  pointer :: prec => null()
    class(abstract_linear_operator_type), pointer :: lin_op => null()
    real(kind=r_def)    :: r_tol
    real(kind=r_def)    :: a_tol
    integer(kind=i_def) :: max_iter
    logical(kind=l_def) :: monitor_convergence
    logical(kind=l_def) :: fail_on_non_converged
  contains
    procedure(apply_interface), deferred :: apply
  end type abstract_iterative_solver_type

  abstract interface
    subroutine apply_interface(self, x, b)
      import :: abstract_vector_type, abstract_iterative_solver_type
      class(abstract_iterative_solver_type), intent(inout) :: self
      class(abstract_vector_type),           intent(inout) :: x
      class(abstract_vector_type),           intent(inout) :: b
    end subroutine apply_interface
  end interface

  type, public, extends(abstract_iterative_solver_type) :: conjugate_gradient_type
    private
    integer(kind=i_def), public :: iterations = 0
  contains
    procedure :: apply => cg_solve
  end type conjugate_gradient_type

  interface conjugate_gradient_type
    module procedure cg_constructor
  end interface

  type :: array_abstract_vector_type
    class(abstract_vector_type), allocatable :: vt
  end type array_abstract_vector_type

  type, public, extends(abstract_iterative_solver_type) :: bicgstab_type
    private
    integer(kind=i_def), public :: iterations = 0
  contains
    procedure :: apply => bicgstab_solve
  end type bicgstab_type

  interface bicgstab_type
    module procedure bicgstab_constructor
  end interface

  type, public, extends(abstract_iterative_solver_type) :: gmres_type
    private
    integer(kind=i_def) :: gcrk
    integer(kind=i_def), public :: cycles = 0
  contains
    procedure :: apply => gmres_solve
  end type gmres_type

  interface gmres_type
    module procedure gmres_constructor
  end interface

contains

  function cg_constructor(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence, &
                          fail_on_non_converged) result(self)
    class(abstract_linear_operator_type), target, intent(in) :: lin_op
    class(abstract_preconditioner_type),  target, intent(in) :: prec
    real(kind=r_def),    intent(in) :: r_tol, a_tol
    integer(kind=i_def), intent(in) :: max_iter
    logical(kind=l_def), intent(in) :: monitor_convergence, fail_on_non_converged
    type(conjugate_gradient_type) :: self
    self%lin_op => lin_op
    self%prec => prec
    self%r_tol = r_tol
    self%a_tol = a_tol
    self%max_iter = max_iter
    self%monitor_convergence = monitor_convergence
    self%fail_on_non_converged = fail_on_non_converged
  end function cg_constructor

  subroutine cg_solve(self, x, b)
    class(conjugate_gradient_type), intent(inout) :: self
    class(abstract_vector_type),    intent(inout) :: x
    class(abstract_vector_type),    intent(inout) :: b
    integer(kind=i_def) :: iter
    real(kind=r_def)    :: alpha, beta, r_nrm, r_nrm_0, rz, rz_new
    logical(kind=l_def) :: converged
    class(abstract_vector_type), allocatable :: r, p, z

    call x%duplicate(r)
    call r%set_scalar(0.0_r_def)
    call x%duplicate(p)
    call p%set_scalar(0.0_r_def)
    call x%duplicate(z)
    converged = .false.
    self%iterations = 0
    call self%lin_op%apply(x, r)
    call r%scale(-1.0_r_def)
    call r%axpy(1.0_r_def, b)
    r_nrm_0 = r%norm()
    if (r_nrm_0 < EPS) return
    call z%set_scalar(0.0_r_def)
    call self%prec%apply(r, z)
    rz = r%dot(z)
    call p%copy(z)
    do iter = 1, self%max_iter
      call self%lin_op%apply(p, z)
      alpha = rz / p%dot(z)
      call x%axpy(alpha, p)
      call r%axpy(-alpha, z)
      self%iterations = iter
      if (self%monitor_convergence) then
        r_nrm = r%norm()
        if (r_nrm / r_nrm_0 <= self%r_tol .or. r_nrm <= self%a_tol) then
          converged = .true.
          exit
        end if
      end if
      call self%prec%apply(r, z)
      rz_new = r%dot(z)
      beta = rz_new / rz
      call p%aypx(beta, z)
      rz = rz_new
    end do
    if (self%monitor_convergence .and. .not. converged) then
      write(log_scratch_space, '("cg failed to converge after ",I6," iterations")') self%max_iter
      if (self%fail_on_non_converged) then
        call log_event(log_scratch_space, LOG_LEVEL_ERROR)
      else
        call log_event(log_scratch_space, LOG_LEVEL_INFO)
      end if
    end if
  end subroutine cg_solve

  function bicgstab_constructor(lin_op, prec, r_tol, a_tol, max_iter, monitor_convergence, &
                                fail_on_non_converged) result(self)
    class(abstract_linear_operator_type), target, intent(in) :: lin_op
    class(abstract_preconditioner_type),  target, intent(in) :: prec
    real(kind=r_def),    intent(in) :: r_tol, a_tol
    integer(kind=i_def), intent(in) :: max_iter
    logical(kind=l_def), intent(in) :: monitor_convergence, fail_on_non_converged
    type(bicgstab_type) :: self
    self%lin_op => lin_op
    self%prec => prec
    self%r_tol = r_tol
    self%a_tol = a_tol
    self%max_iter = max_iter
    self%monitor_convergence = monitor_convergence
    self%fail_on_non_converged = fail_on_non_converged
  end function bicgstab_constructor

  subroutine bicgstab_solve(self, x, b)
    class(bicgstab_type),        intent(inout) :: self
    class(abstract_vector_type), intent(inout) :: x
    class(abstract_vector_type), intent(inout) :: b
    class(abstract_vector_type), allocatable :: r, r0, p, v, t, s, y, z
    real(kind=r_def) :: alpha, beta, rho, omega, rho_old, err, sc_err, tt, ts
    integer(kind=i_def) :: iter

    call x%duplicate(r)
    call r%copy(b)
    call x%duplicate(v)
    call v%set_scalar(0.0_r_def)
    call self%lin_op%apply(x, v)
    call r%axpy(-1.0_r_def, v)
    call r%duplicate(r0)
    call r0%copy(r)
    sc_err = r%norm()
    self%iterations = 0
    if (sc_err < EPS) return
    sc_err = max(sc_err, self%a_tol)
    alpha = 1.0_r_def
    omega = 1.0_r_def
    rho_old = 1.0_r_def
    call x%duplicate(p)
    call x%duplicate(s)
    call x%duplicate(t)
    call v%set_scalar(0.0_r_def)
    call p%set_scalar(0.0_r_def)
    call x%duplicate(y)
    call x%duplicate(z)
    err = huge(1.0_r_def)
    do iter = 1, self%max_iter
      rho = r%dot(r0)
      beta = (rho / rho_old) * (alpha / omega)
      call t%copy(r)
      call t%axpy(-beta * omega, v)
      call self%prec%apply(t, y)
      call p%aypx(beta, y)
      call self%lin_op%apply(p, v)
      alpha = rho / r0%dot(v)
      call s%copy(r)
      call s%axpy(-alpha, v)
      call self%prec%apply(s, z)
      call self%lin_op%apply(z, t)
      tt = t%dot(t)
      ts = t%dot(s)
      omega = ts / tt
      call x%axpy(omega, z)
      call x%axpy(alpha, p)
      call r%copy(s)
      call r%axpy(-omega, t)
      rho_old = rho
      self%iterations = iter
      if (self%monitor_convergence) then
        err = r%norm() / sc_err
        if (err < self%r_tol) exit
      end if
    end do
    if (self%monitor_convergence .and. err >= self%r_tol) then
      write(log_scratch_space, '(A,I4,A,E15.8)') "bicgstab: NOT converged in", self%max_iter, " iters, Res=", err
      if (self%fail_on_non_converged) then
        call log_event(log_scratch_space, LOG_LEVEL_ERROR)
      else
        call log_event(log_scratch_space, LOG_LEVEL_INFO)
      end if
    end if
  end subroutine bicgstab_solve

  function gmres_constructor(lin_op, prec, gcrk, r_tol, a_tol, max_iter, monitor_convergence, &
                             fail_on_non_converged) result(self)
    class(abstract_linear_operator_type), target, intent(in) :: lin_op
    class(abstract_preconditioner_type),  target, intent(in) :: prec
    integer(kind=i_def), intent(in) :: gcrk
    real(kind=r_def),    intent(in) :: r_tol, a_tol
    integer(kind=i_def), intent(in) :: max_iter
    logical(kind=l_def), intent(in) :: monitor_convergence, fail_on_non_converged
    type(gmres_type) :: self
    self%lin_op => lin_op
    self%prec => prec
    self%gcrk = gcrk
    self%r_tol = r_tol
    self%a_tol = a_tol
    self%max_iter = max_iter
    self%monitor_convergence = monitor_convergence
    self%fail_on_non_converged = fail_on_non_converged
  end function gmres_constructor

  subroutine gmres_solve(self, x, b)
    class(gmres_type),           intent(inout) :: self
    class(abstract_vector_type), intent(inout) :: x
    class(abstract_vector_type), intent(inout) :: b
    class(abstract_vector_type), allocatable :: s, w, Ax, res
    type(array_abstract_vector_type), allocatable :: v(:)
    real(kind=r_def), allocatable :: u(:), g(:), h(:,:)
    real(kind=r_def) :: beta, si, ci, nrm, h1, h2, p, q, err, sc_err
    integer(kind=i_def) :: iv, ivj, iter

    call x%duplicate(Ax)
    call Ax%set_scalar(0.0_r_def)
    call b%duplicate(res)
    call self%lin_op%apply(x, Ax)
    call res%copy(b)
    call res%axpy(-1.0_r_def, Ax)
    sc_err = res%norm()
    self%cycles = 0
    if (sc_err < EPS) return
    if (self%monitor_convergence) sc_err = max(sc_err, self%a_tol)
    call x%duplicate(s)
    call x%duplicate(w)
    allocate(v(self%gcrk))
    do iv = 1, self%gcrk
      allocate(v(iv)%vt, source=x)
    end do
    allocate(h(self%gcrk + 1, self%gcrk), g(self%gcrk + 1), u(self%gcrk))
    err = huge(1.0_r_def)
    do iter = 1, self%max_iter
      call s%set_scalar(0.0_r_def)
      call self%prec%apply(res, s)
      beta = s%norm()
      call v(1)%vt%copy(s)
      call v(1)%vt%scale(1.0_r_def / beta)
      g(:) = 0.0_r_def
      g(1) = beta
      do iv = 1, self%gcrk
        call w%copy(v(iv)%vt)
        call self%lin_op%apply(w, s)
        call self%prec%apply(s, w)
        do ivj = 1, iv
          h(ivj, iv) = v(ivj)%vt%dot(w)
          call w%axpy(-h(ivj, iv), v(ivj)%vt)
        end do
        h(iv + 1, iv) = w%norm()
        if (iv < self%gcrk) then
          call v(iv + 1)%vt%copy(w)
          call v(iv + 1)%vt%scale(1.0_r_def / h(iv + 1, iv))
        end if
      end do
      do iv = 1, self%gcrk
        nrm = sqrt(h(iv, iv) * h(iv, iv) + h(iv + 1, iv) * h(iv + 1, iv))
        si = h(iv + 1, iv) / nrm
        ci = h(iv, iv) / nrm
        p = ci * g(iv) + si * g(iv + 1)
        q = -si * g(iv) + ci * g(iv + 1)
        g(iv) = p
        g(iv + 1) = q
        do ivj = iv, self%gcrk
          h1 = ci * h(iv, ivj) + si * h(iv + 1, ivj)
          h2 = -si * h(iv, ivj) + ci * h(iv + 1, ivj)
          h(iv, ivj) = h1
          h(iv + 1, ivj) = h2
        end do
      end do
      u(self%gcrk) = g(self%gcrk) / h(self%gcrk, self%gcrk)
      do iv = self%gcrk - 1, 1, -1
        u(iv) = g(iv)
        do ivj = iv + 1, self%gcrk
          u(iv) = u(iv) - h(iv, ivj) * u(ivj)
        end do
        u(iv) = u(iv) / h(iv, iv)
      end do
      do iv = 1, self%gcrk
        call s%copy(v(iv)%vt)
        call x%axpy(u(iv), s)
      end do
      call Ax%set_scalar(0.0_r_def)
      call self%lin_op%apply(x, Ax)
      call res%copy(Ax)
      call res%aypx(-1.0_r_def, b)
      self%cycles = iter
      if (self%monitor_convergence) then
        err = res%norm() / sc_err
        if (err < self%r_tol) exit
      end if
    end do
    if (ieee_is_nan(err) .or. (self%monitor_convergence .and. err >= self%r_tol .and. &
                               self%fail_on_non_converged)) then
      call log_event("GMRES solver_algorithm: NOT converged", LOG_LEVEL_ERROR)
    end if
    deallocate(h, g, u)
    do iv = 1, self%gcrk
      deallocate(v(iv)%vt)
    end do
    deallocate(v)
  end subroutine gmres_solve

end module sci_iterative_solver_mod
