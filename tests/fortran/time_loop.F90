! Test driver: the time loop of main.f95:247-322 (velnw, bondv1 or
! bondv1_data24, velfg, feedbf, les, adam, press, and aveflow, with the
! reference's argument lists), compiled
! against the drop-in modules of mpi-les_amd/fortran/les_amd_modules.F90,
! behind one bind(C) entry point so a test can hand it seeded arrays.
!
! The arrays main.f95 declares and this loop only passes along (diu, cov,
! nou, dfu1/dfv1/dfw1, fx/fy/fz) are local here.  On return the state arrays
! hold the device state (les_sync) and every host array is forgotten by the
! library (les_release_all), so the locals can be freed.  data20/data21 are
! written (n1 == nmax) into files named by the caller.  With jdata24 /= 0
! the inflow comes from data24 records in the directory dir24
! (main.f95:264-268); with nave1 > 0 aveflow runs after every step, averaging
! from step nave1, and writes data10/data11 at n == nmax (main.f95:320-322).
module les_time_loop_test
    use iso_c_binding
    implicit none
contains
    subroutine time_loop(im, jm, km, n0, n1, nmax, ro, dt, vn, alpha, beta, &
            dx1, dy1, dzn, dzs, dxs, dys, z2, cn1, cn2l, cn2s, cn3l, cn3s, cn4l, cn4s, &
            bmask1, cmask1, dmask1, p, u, v, w, f, g, h, usum, vsum, wsum, fold, gold, hold, fghold, &
            rhs, sm, uspd, vspd, delx1, c_data20, c_data21, jdata24, c_dir24, nave1, aveu, avev, avew, &
            avep, avel, aveuu, avevv, aveww, avesm, avesmsm, uwfx, c_data10, c_data11) &
            bind(C, name='les_test_time_loop')
        use les_amd_c
        use module_velnw
        use module_bondv1
        use module_velFG
        use module_feedbf
        use module_les
        use module_adam
        use module_press
        use module_bondv1_data24
        use module_aveflow
        integer(c_int), value :: im, jm, km, n0, n1, nmax, jdata24, nave1
        real(c_float), value :: ro, dt, vn, alpha, beta
        real(c_float), dimension(-1:im+1) :: dx1
        real(c_float), dimension(0:jm+1) :: dy1
        real(c_float), dimension(-1:km+2) :: dzn, dzs
        real(c_float), dimension(0:im) :: dxs
        real(c_float), dimension(0:jm) :: dys
        real(c_float), dimension(0:km+2) :: z2
        real(c_float), dimension(im,jm,km) :: cn1, fold, gold, hold, fghold
        real(c_float), dimension(im) :: cn2l, cn2s
        real(c_float), dimension(jm) :: cn3l, cn3s
        real(c_float), dimension(km) :: cn4l, cn4s, delx1
        real(c_float), dimension(-1:im+1,0:jm+1,0:km+1) :: bmask1
        real(c_float), dimension(0:im+1,-1:jm+1,0:km+1) :: cmask1, u, v
        real(c_float), dimension(0:im+1,0:jm+1,0:km+1) :: dmask1, rhs
        real(c_float), dimension(0:im+2,0:jm+2,0:km+1) :: p
        real(c_float), dimension(0:im+1,-1:jm+1,-1:km+1) :: w
        real(c_float), dimension(0:im,0:jm,0:km) :: f, g, h, usum, vsum, wsum
        real(c_float), dimension(-1:im+1,-1:jm+1,0:km+1) :: sm
        real(c_float), dimension(0:im+1,0:jm+1) :: uspd, vspd
        character(kind=c_char), dimension(70) :: c_data20, c_data21, c_data10, c_data11
        character(kind=c_char), dimension(512) :: c_dir24
        real(c_float), dimension(im,jm,0:km) :: aveu, avev
        real(c_float), dimension(im+1,jm,0:km+2) :: avew
        real(c_float), dimension(im,jm,km) :: avep, avel, aveuu, avevv, aveww, avesm, avesmsm, uwfx
        real(c_float), allocatable, dimension(:,:) :: avesu, avesv, avesw, avesuu, avesvv, avesww, uwfxs
        real(c_float), allocatable, dimension(:,:,:) :: amask1
        character(len=70) :: data10, data11, data13, data14
        real(c_float) :: time

        real(c_float), allocatable, dimension(:,:,:) :: diu1, diu2, diu3, diu4, diu5, diu6, diu7, &
            diu8, diu9, cov1, cov2, cov3, cov4, cov5, cov6, cov7, cov8, cov9, nou1, nou2, nou3, nou4, &
            nou5, nou6, nou7, nou8, nou9, dfu1, dfv1, dfw1, fx, fy, fz
        character(len=70) :: data20, data21
        integer :: n, i

        do i = 1, 70
            data20(i:i) = c_data20(i)
            data21(i:i) = c_data21(i)
            data10(i:i) = c_data10(i)
            data11(i:i) = c_data11(i)
        end do
        data13 = ''
        data14 = ''
        do i = 1, 512
            data24_dir(i:i) = c_dir24(i)
        end do
        data24_dir_known = .true.
        allocate(avesu(im,km), avesv(im,km), avesw(im,km), avesuu(im,km), avesvv(im,km), avesww(im,km), &
                 uwfxs(im,km), amask1(0:im+1,0:jm+1,0:km+1))
        amask1 = 0.
        time = 0.
        allocate(diu1(-1:im+2,0:jm+2,0:km+2), diu5(-1:im+2,0:jm+2,0:km+2), &
                 cov1(-1:im+2,0:jm+2,0:km+2), cov5(-1:im+2,0:jm+2,0:km+2), &
                 nou1(-1:im+2,0:jm+2,0:km+2), nou5(-1:im+2,0:jm+2,0:km+2))
        allocate(diu2(0:im+2,0:jm+2,0:km+2), diu3(0:im+2,0:jm+2,0:km+2), diu4(0:im+2,0:jm+2,0:km+2), &
                 diu6(0:im+2,0:jm+2,0:km+2), diu7(0:im+2,0:jm+2,0:km+2), diu8(0:im+2,0:jm+2,0:km+2), &
                 diu9(0:im+2,0:jm+2,0:km+2))
        allocate(cov2(0:im+2,0:jm+2,0:km+2), cov3(0:im+2,0:jm+2,0:km+2), cov4(0:im+2,0:jm+2,0:km+2), &
                 cov6(0:im+2,0:jm+2,0:km+2), cov7(0:im+2,0:jm+2,0:km+2), cov8(0:im+2,0:jm+2,0:km+2), &
                 cov9(0:im+2,0:jm+2,0:km+2))
        allocate(nou2(0:im+2,0:jm+2,0:km+2), nou3(0:im+2,0:jm+2,0:km+2), nou4(0:im+2,0:jm+2,0:km+2), &
                 nou6(0:im+2,0:jm+2,0:km+2), nou7(0:im+2,0:jm+2,0:km+2), nou8(0:im+2,0:jm+2,0:km+2), &
                 nou9(0:im+2,0:jm+2,0:km+2))
        allocate(dfu1(0:im,jm,km), dfv1(im,0:jm,km), dfw1(im,jm,km))
        allocate(fx(0:im,0:jm,0:km), fy(0:im,0:jm,0:km), fz(0:im,0:jm,0:km))
        diu1 = 0.; diu2 = 0.; diu3 = 0.; diu4 = 0.; diu5 = 0.; diu6 = 0.; diu7 = 0.; diu8 = 0.; diu9 = 0.

        do n = n0, n1
            call velnw(km,jm,im,p,ro,dxs,u,dt,f,dys,v,g,dzs,w,h)
            if (jdata24 == 0) then
                call bondv1(jm,u,z2,dzn,v,w,km,n,im,dt,dxs)
            else
                call bondv1_data24(jm,u,z2,dzn,v,w,km,n,im,dt,dxs)
            end if
            call velfg(km,jm,im,dx1,cov1,cov2,cov3,dfu1,diu1,diu2,dy1,diu3,dzn, &
                       vn,f,cov4,cov5,cov6,dfv1,diu4,diu5,diu6,g,cov7,cov8,cov9, &
                       dfw1,diu7,diu8,diu9,dzs,h,nou1,u,nou5,v,nou9,w,nou2,nou3, &
                       nou4,nou6,nou7,nou8,uspd,vspd)
            call feedbf(km,jm,im,usum,u,bmask1,vsum,v,cmask1,wsum,w,dmask1,alpha, &
                        dt,beta,fx,fy,fz,f,g,h)
            call les(km,delx1,dx1,dy1,dzn,jm,im,diu1,diu2,diu3,diu4,diu5,diu6, &
                     diu7,diu8,diu9,sm,f,g,h,uspd,vspd,dxs,dys)
            call adam(n,nmax,data21,fold,im,jm,km,gold,hold,fghold,f,g,h)
            call press(km,jm,im,rhs,u,dx1,v,dy1,w,dzn,f,g,h,dt,cn1,cn2l,p,cn2s, &
                       cn3l,cn3s,cn4l,cn4s,n, nmax,data20,usum,vsum,wsum)
            if (nave1 > 0) then
                call aveflow(n,nave1,km,jm,im,aveu,avev,avew,avep,avel,aveuu,avevv,aveww, &
                             avesm,avesmsm,uwfx,avesu,avesv,avesw,avesuu,avesvv, &
                             avesww,u,v,w,p,sm,nmax,uwfxs,data10,time,data11,data13,data14,amask1)
            end if
        end do

        call les_sync(p); call les_sync(u); call les_sync(v); call les_sync(w)
        call les_sync(f); call les_sync(g); call les_sync(h)
        call les_sync(usum); call les_sync(vsum); call les_sync(wsum)
        call les_sync(fold); call les_sync(gold); call les_sync(hold)
        call les_sync(rhs); call les_sync(sm); call les_sync(uspd); call les_sync(vspd)
        call les_check(les_release_all(), 'release_all')
    end subroutine time_loop
end module les_time_loop_test
